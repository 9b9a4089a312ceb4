// dme_kernels.hip -- gfx950 kernels of the Dirac motion-estimation path.
//
// Layout: every picture plane is int16, row-major, stride = width, in HBM.
// MEData arrays are the reference's (motion.h) flattened row-major.  One
// launch serves every picture of a batch: work items are mapped to
// (job, reference) tasks through a sorted offset table.
//
// Work split.  The reference is a raster-order sequential search; its only
// cross-block dependencies are the vector predictions (and, in mode
// decision, superblock merging).  Every candidate SAD that does not depend on
// a prediction is computed in bulk, fully parallel (k_pel_bulk, k_mode_l2);
// the dependent remainder runs as an anti-diagonal wavefront inside one
// workgroup per (picture, reference): t = x + 2y for blocks
// (pixel_match.cpp:314-347, me_subpel.cpp:131-176), t = X + Y for
// superblocks (me_mode_decn.cpp:161-165, 308-321).  A wavefront step only
// reads results of earlier steps, separated by a workgroup barrier.
//
// Inside a wave, candidate SADs are computed by groups of R lanes per
// candidate (R rows each), reduced with xor shuffles, and the winner is the
// minimum of the 64-bit key (SAD << 32 | enumeration index): the first
// minimum in the reference's enumeration order, i.e. its strict-< argmin.
//
// Float costs follow the reference bit for bit: no contraction (built with
// -ffp-contract=off), the same float/double types and operation order.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <string>
#include <type_traits>

#include "dme_device.h"
#include "dme_internal.h"

namespace dme {

#define WAVE 64

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Workgroups are dispatched to the 8 XCDs round-robin, and each XCD has its
// own L2.  xcd_block maps the dispatch index to a logical index so that each
// XCD receives one contiguous range of the grid: neighbouring blocks, whose
// reference regions overlap, then share an L2.
__device__ __forceinline__ int xcd_block() {
    const int n = gridDim.x, b = blockIdx.x;
    const int q = n >> 3, r = n & 7, x = b & 7, i = b >> 3;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// Table kernels: consecutive blocks (or units) per wave, so that a launch
// has fewer, longer workgroups and a wave's blocks share reference data
// (measured per kernel; the pel and sub-pel tables keep one block per wave:
// a loop there costs registers and occupancy).
#define PW_MODE_L2 2
#define PW_MODE_TABLES 4

// Index of the calling wave in its workgroup, known to the compiler as
// wave-uniform, so that what is derived from it stays in scalar registers.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// The task of work item gi for a wave whose lanes all hold the same gi:
// the number of tasks starting at or before gi, counted by a ballot over
// one coalesced load per 64 tasks (one memory round trip instead of a
// binary search's chain of dependent loads).  Tasks are sorted by offset.
__device__ __forceinline__ int find_task_wave(const Task *t, int n, int gi) {
    const int lane = threadIdx.x & 63;
    int cnt = 0;
    for (int j0 = 0; j0 < n; j0 += 64) {
        const int j = j0 + lane;
        cnt += __popcll(__ballot(j < n && t[j].offset <= gi));
    }
    return cnt - 1;
}

__device__ __forceinline__ int find_task(const Task *t, int n, int gi) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t[mid].offset <= gi) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
    for (int m = 32; m >= 1; m >>= 1) {
        const unsigned long long o = __shfl_xor(v, m, WAVE);
        v = o < v ? o : v;
    }
    return v;
}

// Sum over the 64 lanes of the wave (all active), returned to every lane:
// an inclusive scan within each row of 16 lanes by DPP row shifts (1, 2, 4,
// 8; a lane whose source is outside its row adds 0), the rows joined by the
// row broadcasts of lanes 15 and 31, the total read from lane 63.
__device__ __forceinline__ int wave_sum(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 to rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 to rows 2, 3
    return __builtin_amdgcn_readlane(v, 63);
}

// BlockDiffParams::SetBlockLimits (me_utils.cpp:58-78)
struct Blk {
    int xp, yp, xl, yl;
    __device__ bool empty() const { return xl <= 0 || yl <= 0; }
};

__device__ __forceinline__ Blk block_limits(const int blen[2], const int bsep[2], const int boff[2], int W, int H,
                                            int xb, int yb) {
    const int lx = xb * bsep[0] - boff[0], ly = yb * bsep[1] - boff[1];
    Blk b;
    b.xp = lx > 0 ? lx : 0;
    b.yp = ly > 0 ? ly : 0;
    b.xl = blen[0] - b.xp + lx;
    b.yl = blen[1] - b.yp + ly;
    if (b.xp + b.xl - 1 > W - 1) b.xl = W - b.xp;
    if (b.yp + b.yl - 1 > H - 1) b.yl = H - b.yp;
    return b;
}

__device__ __forceinline__ int norm1(int x, int y) { return abs(x) + abs(y); }

// MvMedian of three (motion.cpp:500-515), MvMean (motion.cpp:640-651)
__device__ __forceinline__ dme_mv med3(dme_mv a, dme_mv b, dme_mv c) {
    dme_mv r;
    r.x = a.x + b.x + c.x - max(max(a.x, b.x), c.x) - min(min(a.x, b.x), c.x);
    r.y = a.y + b.y + c.y - max(max(a.y, b.y), c.y) - min(min(a.y, b.y), c.y);
    return r;
}
__device__ __forceinline__ dme_mv mean2(dme_mv a, dme_mv b) {
    dme_mv r;
    r.x = (a.x + b.x + 1) >> 1;
    r.y = (a.y + b.y + 1) >> 1;
    return r;
}

// Window enumeration of AddNewVlist (block_match.cpp:89-126): index k of the
// (2r+1)^2 square around the centre.
__device__ __forceinline__ dme_mv win_off(int k, int r) {
    dme_mv m;
    if (k == 0) { m.x = 0; m.y = 0; return m; }
    if (k <= 2 * r) {
        const int i = (k + 1) >> 1;
        m.x = (k & 1) ? i : -i;
        m.y = 0;
        return m;
    }
    const int q = k - 2 * r - 1, w = 2 * (2 * r + 1);
    const int j = q / w + 1, rem = q - (j - 1) * w;
    m.x = (rem >> 1) - r;
    m.y = (rem & 1) ? -j : j;
    return m;
}

// Lanes per candidate: the power of two minimising passes x rows per lane.
__device__ __forceinline__ int pick_group(int nc, int rows) {
    int best_r = 1, best_w = INT_MAX;
    for (int r = 1; r <= WAVE; r <<= 1) {
        const int per = WAVE / r;
        const int w = ((nc + per - 1) / per) * ((rows + r - 1) / r);
        if (w < best_w) { best_w = w; best_r = r; }
    }
    return best_r;
}

// ---------------------------------------------------------------------------
// Integer pel SADs (PelBlockDiff, me_utils.cpp:138-282): clamp-to-edge ref
// reads; equal to both reference paths.  Returns the wave-uniform key of the
// first minimum over candidates 0..nc-1 produced by cand(c).
// ---------------------------------------------------------------------------
template <class CandFn>
__device__ unsigned long long wave_pel_argmin(const int16_t *__restrict__ pic, const int16_t *__restrict__ ref,
                                              int W, int H, const Blk &b, int nc, CandFn cand) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int R = pick_group(nc, b.yl);
    const int per = WAVE / R, g = lane / R, sub = lane - g * R;
    unsigned long long best = ~0ull;
    for (int base = 0; base < nc; base += per) {
        const int c = base + g;
        int s = 0;
        if (c < nc) {
            const dme_mv mv = cand(c);
            const bool inx = b.xp + mv.x >= 0 && b.xp + b.xl + mv.x <= W;
            for (int row = sub; row < b.yl; row += R) {
                const int y = b.yp + row;
                const int16_t *pr = pic + (size_t)y * W + b.xp;
                const int16_t *rr = ref + (size_t)clampi(y + mv.y, 0, H - 1) * W;
                if (inx) {
                    const int16_t *rq = rr + b.xp + mv.x;
                    for (int x = 0; x < b.xl; ++x) s += abs((int)pr[x] - (int)rq[x]);
                } else {
                    for (int x = 0; x < b.xl; ++x) s += abs((int)pr[x] - (int)rr[clampi(b.xp + x + mv.x, 0, W - 1)]);
                }
            }
        }
        for (int m = 1; m < R; m <<= 1) s += __shfl_xor(s, m, WAVE);
        unsigned long long key = (c < nc && sub == 0) ? (((unsigned long long)(unsigned)s << 32) | (unsigned)c) : ~0ull;
        key = wave_min_u64(key);
        best = key < best ? key : best;
    }
    return best;
}

// ---------------------------------------------------------------------------
// K1: DownConverter::DoDownConvert + RowLoop (downconvert.cpp:51-204).
// Separable: the vertical 12-tap pass fills the short row buffer, then the
// horizontal pass filters it; both round (+128) >> 8 and the row buffer is
// short.  Index clamping to [0, 2*len-1] is the reference's edge handling for
// len >= 12 (host-checked).  A workgroup makes a tile of 64 x 8 output
// samples from a 138 x 26 input tile staged in LDS.
// ---------------------------------------------------------------------------
__constant__ int c_down_taps[6] = {86, 46, 4, -8, -4, 4};

#define DC_TW kDownTileW          // output tile width
#define DC_TH kDownTileH          // output tile height
#define DC_IW (2 * DC_TW + 10)    // input tile columns (138)
#define DC_IH (2 * DC_TH + 10)    // input tile rows (42)

// One launch serves every plane of a batch: tiles are mapped to planes
// through the offset table.
__global__ __launch_bounds__(256) void k_downconvert(const PlaneTask *__restrict__ pt, int npt, int total) {
    __shared__ int16_t s_in[DC_IH][DC_IW];
    __shared__ int16_t s_v[DC_TH][DC_IW];
    const int tile = blockIdx.x;
    if (tile >= total) return;
    int lo = 0, hi = npt - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pt[mid].offset <= tile) lo = mid;
        else hi = mid - 1;
    }
    const PlaneTask p = pt[lo];
    const int local = tile - p.offset;
    const int tx = (p.ow + DC_TW - 1) / DC_TW;
    const int cx0 = (local % tx) * DC_TW, cy0 = (local / tx) * DC_TH;
    const int ow = p.ow, oh = p.oh, in_w = p.in_w;
    const int xlen = 2 * ow, ylen = 2 * oh;
    const int16_t *__restrict__ in = p.in;
    int16_t *__restrict__ out = p.out;
    const int tid = threadIdx.x;
    // input tile: rows 2*cy0 - 5 + r, columns 2*cx0 - 5 + c, clamped
    for (int i0 = 0; i0 < DC_IH * DC_IW; i0 += 8 * 256) {
        int16_t v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int i = i0 + k * 256 + tid;
            const int r = i / DC_IW, c = i - r * DC_IW;
            v[k] = 0;
            if (i < DC_IH * DC_IW)
                v[k] = in[(size_t)clampi(2 * cy0 - 5 + r, 0, ylen - 1) * in_w + clampi(2 * cx0 - 5 + c, 0, xlen - 1)];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int i = i0 + k * 256 + tid;
            if (i < DC_IH * DC_IW) (&s_in[0][0])[i] = v[k];
        }
    }
    __syncthreads();
    // vertical pass: output row o uses input rows 2o .. 2o + 11 of the tile
    for (int i = tid; i < DC_TH * DC_IW; i += 256) {
        const int o = i / DC_IW, c = i - o * DC_IW;
        int sum = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) sum += ((int)s_in[2 * o + 5 - k][c] + (int)s_in[2 * o + 6 + k][c]) * c_down_taps[k];
        s_v[o][c] = (int16_t)((sum + 128) >> 8);
    }
    __syncthreads();
    for (int i = tid; i < DC_TH * DC_TW; i += 256) {
        const int o = i / DC_TW, c = i - o * DC_TW;
        const int cx = cx0 + c, cy = cy0 + o;
        if (cx >= ow || cy >= oh) continue;
        int sum = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) sum += ((int)s_v[o][2 * c + 5 - k] + (int)s_v[o][2 * c + 6 + k]) * c_down_taps[k];
        out[(size_t)cy * ow + cx] = (int16_t)((sum + 128) >> 8);
    }
}

// ---------------------------------------------------------------------------
// K2: UpConverter::DoUpConverter (upconvert.cpp:55-200), clip [-128,127].
// The accumulator is a short in the reference; |sum| <= 16 + 32*255 fits, so
// int arithmetic gives the same value before the (short) >> 5.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int up_tap8(const int s[8]) {
    const int sum = 16 + (s[3] + s[4]) * 21 + (s[2] + s[5]) * -7 + (s[1] + s[6]) * 3 + (s[0] + s[7]) * -1;
    return clampi(sum >> 5, -128, 127);
}

// A workgroup filters a 64 x 16 tile of source samples: the tile and its
// apron (4 columns left, 4 right, 3 rows above, 4 below, clamped to the
// picture) are staged in LDS once, the vertical filter of every staged column
// goes to LDS (4 rows per thread, the filter sliding down the column), and
// each thread then writes the 2 x 2 quads of 4 consecutive columns of one
// row (the filter sliding along it), 16 bytes per up row.  U8: the plane of
// motion compensation, one byte per sample biased by 128 (every sample of a
// picture in 8-bit range lies in [-128, 127]: the source samples and the
// clipped taps), 8 bytes per up row.
constexpr int UC_IW = kUpTileW + 8, UC_IH = kUpTileH + 7;  // columns x0 - 4 .. x0 + 67

__device__ __forceinline__ int up_tap8_lds(const int16_t *r) {
    const int sum = 16 + (r[3] + r[4]) * 21 + (r[2] + r[5]) * -7 + (r[1] + r[6]) * 3 + (r[0] + r[7]) * -1;
    return clampi(sum >> 5, -128, 127);
}

__device__ __forceinline__ uint32_t pack16(int lo, int hi) { return (uint32_t)(uint16_t)lo | ((uint32_t)(uint16_t)hi << 16); }

__device__ __forceinline__ uint32_t pack8x2(int a, int b, int c, int d) {
    return (uint32_t)(a + 128) | ((uint32_t)(b + 128) << 8) | ((uint32_t)(c + 128) << 16) | ((uint32_t)(d + 128) << 24);
}

template <bool U8>
__global__ __launch_bounds__(256) void k_upconvert(const PlaneTask *__restrict__ pt, int npt, int total) {
    __shared__ int16_t s_in[UC_IH][UC_IW + 1];
    __shared__ int16_t s_vt[kUpTileH][UC_IW + 1];
    const int tile = blockIdx.x;
    if (tile >= total) return;
    int lo = 0, hi = npt - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pt[mid].offset <= tile) lo = mid;
        else hi = mid - 1;
    }
    const PlaneTask p = pt[lo];
    const int local = tile - p.offset;
    const int W = p.ow, H = p.oh;
    const int tx = (W + kUpTileW - 1) / kUpTileW;
    const int x0t = (local % tx) * kUpTileW, y0 = (local / tx) * kUpTileH;
    const int16_t *__restrict__ in = p.in;
    const int tid = threadIdx.x;
    for (int i = tid; i < UC_IH * UC_IW; i += 256) {
        const int r = i / UC_IW, c = i - r * UC_IW;
        s_in[r][c] = in[(size_t)clampi(y0 - 3 + r, 0, H - 1) * W + clampi(x0t - 4 + c, 0, W - 1)];
    }
    __syncthreads();
    // vertical taps: a thread makes 4 consecutive rows of one staged column
    for (int i = tid; i < (kUpTileH / 4) * UC_IW; i += 256) {
        const int g = i / UC_IW, c = i - g * UC_IW, r0 = 4 * g;
        int16_t col[11];
#pragma unroll
        for (int t = 0; t < 11; ++t) col[t] = s_in[r0 + t][c];
#pragma unroll
        for (int k = 0; k < 4; ++k) s_vt[r0 + k][c] = (int16_t)up_tap8_lds(&col[k]);
    }
    __syncthreads();
    // a thread writes 4 consecutive columns of one source row: the row's 11
    // staged values (and the vertical taps' 11) are read once
    const int row = tid >> 4, x0 = x0t + 4 * (tid & 15);
    const int y = y0 + row;
    if (y >= H || x0 >= W) return;
    int16_t wi[11], wv[11];
#pragma unroll
    for (int t = 0; t < 11; ++t) {
        wi[t] = s_in[row + 3][x0 - x0t + 1 + t];  // columns x0 - 3 .. x0 + 7
        wv[t] = s_vt[row][x0 - x0t + 1 + t];
    }
    if (U8) {
        // up row 2y: sample, tap, sample, tap ... of columns x0 .. x0 + 3; 2W bytes per up row
        uint8_t *__restrict__ up8 = reinterpret_cast<uint8_t *>(p.out);
        uint8_t *r0 = up8 + (size_t)(2 * y) * (2 * W) + 2 * x0, *r1 = r0 + 2 * W;
        int ta[4], tb[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            ta[k] = up_tap8_lds(&wi[k]);
            tb[k] = up_tap8_lds(&wv[k]);
        }
        if ((W & 3) == 0) {  // rows start 8-byte aligned: one 8-byte store per row
            *reinterpret_cast<uint2 *>(r0) = make_uint2(pack8x2(wi[3], ta[0], wi[4], ta[1]), pack8x2(wi[5], ta[2], wi[6], ta[3]));
            *reinterpret_cast<uint2 *>(r1) = make_uint2(pack8x2(wv[3], tb[0], wv[4], tb[1]), pack8x2(wv[5], tb[2], wv[6], tb[3]));
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (x0 + k >= W) break;
                r0[2 * k] = (uint8_t)(wi[k + 3] + 128);
                r0[2 * k + 1] = (uint8_t)(ta[k] + 128);
                r1[2 * k] = (uint8_t)(wv[k + 3] + 128);
                r1[2 * k + 1] = (uint8_t)(tb[k] + 128);
            }
        }
        return;
    }
    uint32_t *__restrict__ up = reinterpret_cast<uint32_t *>(p.out);  // two samples per dword, W per up row
    uint32_t *u0 = up + (size_t)(2 * y) * W, *u1 = up + (size_t)(2 * y + 1) * W;
    uint32_t a[4], b[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        a[k] = pack16(wi[k + 3], up_tap8_lds(&wi[k]));
        b[k] = pack16(wv[k + 3], up_tap8_lds(&wv[k]));
    }
    if ((W & 3) == 0) {  // rows start 16-byte aligned: one 16-byte store per row
        *reinterpret_cast<uint4 *>(u0 + x0) = make_uint4(a[0], a[1], a[2], a[3]);
        *reinterpret_cast<uint4 *>(u1 + x0) = make_uint4(b[0], b[1], b[2], b[3]);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (x0 + k >= W) break;
            u0[x0 + k] = a[k];
            u1[x0 + k] = b[k];
        }
    }
}

// ---------------------------------------------------------------------------
// Wavefront order.  The wavefront kernels visit block (x, y) of an X x Y grid
// at step t = x + 2y, so the bulk kernels store each block's record at its
// rank in the order (t, y): the records a step needs are then one contiguous
// run, which the workgroup copies into LDS with coalesced loads.
// ---------------------------------------------------------------------------
// number of blocks with x + 2y < t: the sum over rows of clamp(t - 2y, 0, X)
__host__ __device__ __forceinline__ int wf_before(int t, int X, int Y) {
    if (t <= 0) return 0;
    const int a = t >= X ? min((t - X) / 2 + 1, Y) : 0;  // rows entirely before t
    const int c = min((t + 1) / 2, Y);                   // rows with some block before t
    const int m = c - a;
    return a * X + m * t - m * (a + c - 1);
}
// first row that has a block on diagonal t
__host__ __device__ __forceinline__ int wf_ylo(int t, int X) { return t - X + 1 > 0 ? (t - X + 2) / 2 : 0; }
__host__ __device__ __forceinline__ int wf_pos(int x, int y, int X, int Y) {
    const int t = x + 2 * y;
    return wf_before(t, X, Y) + y - wf_ylo(t, X);
}

// Workgroup copy of diagonal tn's records (rec ints each) from src into the
// LDS slots of bank tn & 1: the first PF loads per thread are issued by
// wf_issue before a step's work and stored by wf_finish after it.
template <int PF>
struct WfCopy {
    int v[PF];
    int len;
    const int *src;
    int *dst;
};

template <int PF>
__device__ __forceinline__ void wf_issue(WfCopy<PF> &c, const int *__restrict__ tab, int *slots, int rec, int max_rows,
                                         int tn, int X, int Y) {
    const int r0 = wf_before(tn, X, Y);
    c.len = (wf_before(tn + 1, X, Y) - r0) * rec;
    c.src = tab + (size_t)r0 * rec;
    c.dst = slots + ((tn & 1) * max_rows + wf_ylo(tn, X)) * rec;
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const int i = k * blockDim.x + threadIdx.x;
        c.v[k] = i < c.len ? c.src[i] : 0;
    }
}

template <int PF>
__device__ __forceinline__ void wf_finish(WfCopy<PF> &c) {
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const int i = k * blockDim.x + threadIdx.x;
        if (i < c.len) c.dst[i] = c.v[k];
    }
    for (int b0 = PF * blockDim.x; b0 < c.len; b0 += PF * blockDim.x) {
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const int i = b0 + k * blockDim.x + threadIdx.x;
            c.v[k] = i < c.len ? c.src[i] : 0;
        }
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const int i = b0 + k * blockDim.x + threadIdx.x;
            if (i < c.len) c.dst[i] = c.v[k];
        }
    }
}

// Luma samples are 8-bit values minus 128 (pic_io.cpp:396-401); the packed
// byte SADs rely on that range, so pictures are checked when registered.
__global__ __launch_bounds__(256) void k_check_range(const int16_t *__restrict__ y, size_t n, int *flag) {
    bool bad = false;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        bad = bad || y[i] < -128 || y[i] > 127;
    if (bad) *flag = 1;
}

// ---------------------------------------------------------------------------
// Combined-component pictures (EncPicture::Combine, enc_picture.cpp:285-367,
// 4:2:0): c = ValueType(sqrt(u^2 + v^2 + (y + 128)^2) - 128), truncated
// toward zero, for each luma sample with the chroma sample of its 2 x 2
// group.  With 8-bit samples x = u^2 + v^2 + (y + 128)^2 is an integer below
// 2^17, exact in the reference's float squares and sums, and its correctly
// rounded float square root never rounds onto or off an integer (an integer
// n <= 313 is more than 2^-15 away from sqrt(n^2 +- 1)).  So with
// s = floor(sqrt(x)): c = s - 128 when x is a square or s >= 128, else
// s - 127 (the negative non-integer rounds toward zero).  Integer work only.
// One thread per sample; CombTask.offset counts 256-sample tiles.
__global__ __launch_bounds__(256) void k_combine(const CombTask *__restrict__ ct, int nct, int total) {
    const int tile = blockIdx.x;
    if (tile >= total) return;
    int lo = 0, hi = nct - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ct[mid].offset <= tile) lo = mid;
        else hi = mid - 1;
    }
    const CombTask c = ct[lo];
    const size_t i = (size_t)(tile - c.offset) * 256 + threadIdx.x;
    if (i >= (size_t)c.W * c.H) return;
    const int y = (int)(i / c.W), x = (int)(i - (size_t)y * c.W);
    const size_t k = (size_t)(y >> 1) * (c.W >> 1) + (x >> 1);
    const int yv = c.y[i] + 128, uv = c.u[k], vv = c.v[k];
    const int q = uv * uv + vv * vv + yv * yv;
    int r = (int)sqrtf((float)q);
    while (r * r > q) --r;
    while ((r + 1) * (r + 1) <= q) ++r;
    c.out[i] = (int16_t)(r * r == q || r >= 128 ? r - 128 : r - 127);
}

// ---------------------------------------------------------------------------
// Quarter-pel planes.  Every value a sub-pel Diff can read from the
// up-converted reference (BlockDiffHalfPel / QuarterPel, me_utils.cpp:410-867)
// is the sample at a quarter-pel position (4X + fx, 4Y + fy): the rounded
// average of the one, two or four up-plane samples around it, with clamped
// coordinates.  They are tabulated once per reference as biased bytes
// (v + 128; up-plane values lie in [-128, 127]) in the layout
//   qp[((Y + 1) * 4 + fy) * (W + 2) + X + 1] = dword of the four fx bytes,
// for X in [-1, W], Y in [-1, H]; positions further out repeat the border
// (both clamped corners coincide there), so reads clamp X and Y to that range.
// A half-pel vector m reads the even phases: position 2m in quarter pels.
// ---------------------------------------------------------------------------
// Built straight from the picture: a workgroup takes 64 x 16 positions (a
// thread per column, four rows apart); the up-plane samples they use (rows 2Y .. 2Y + 2, columns
// 2X .. 2X + 2, clamped) are made in LDS first, as K2 makes them: the
// clamped source tile, its vertical taps, then the 9 x 129 up samples.
#define QT_X 64
#define QT_Y 16
#define QT_SR (QT_Y + 9)       // source rows Y0 - 4 .. Y0 + QT_Y + 4
#define QT_SC (QT_X + 9)       // source columns X0 - 4 .. X0 + QT_X + 4
#define QT_VR (QT_Y + 2)       // vertical-tap rows Y0 - 1 .. Y0 + QT_Y
#define QT_UR (2 * QT_Y + 1)   // up rows 2Y0 .. 2Y0 + 2 QT_Y
#define QT_UC (2 * QT_X + 1)   // up columns 2X0 .. 2X0 + 2 QT_X
// UP8: the same tiles make the byte up-plane instead (up8_* below): biased
// up samples v + 128 of up rows and columns -64 .. 2 H + 63 / 2 W + 63 (a
// replicated margin of 64: the reference's clamped reads for those), rows of
// up8_pitch(W) bytes.
template <bool UP8>
__global__ __launch_bounds__(256) void k_build_qp(const PlaneTask *__restrict__ pt, int npt, int total) {
    __shared__ int16_t s_src[QT_SR][QT_SC];
    __shared__ int16_t s_v[QT_VR][QT_SC];
    __shared__ int16_t s_u[QT_UR][QT_UC + 1];
    const int tile = blockIdx.x;
    if (tile >= total) return;
    int lo = 0, hi = npt - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pt[mid].offset <= tile) lo = mid;
        else hi = mid - 1;
    }
    const PlaneTask p = pt[lo];
    const int W = p.ow, H = p.oh, QW = W + 2, QH = H + 2;
    const int tx = UP8 ? up8_tiles_x(W) : (QW + QT_X - 1) / QT_X, local = tile - p.offset;
    // first position of the tile
    const int X0 = (local % tx) * QT_X - (UP8 ? kUp8Margin / 2 : 1);
    const int Y0 = (local / tx) * QT_Y - (UP8 ? kUp8Margin / 2 : 1);
    // origin of the source tile: the clamped coordinates of every position of
    // the tile lie in SX0 .. SX0 + QT_X, SY0 .. SY0 + QT_Y (for UP8 tiles in
    // the margin the tile itself can lie outside the picture)
    const int SX0 = UP8 ? clampi(X0, 0, max(W - QT_X, 0)) : X0;
    const int SY0 = UP8 ? clampi(Y0, 0, max(H - QT_Y, 0)) : Y0;
    const int16_t *__restrict__ in = p.in;
    const int tid = threadIdx.x;
    for (int i = tid; i < QT_SR * QT_SC; i += 256) {
        const int r = i / QT_SC, c = i - r * QT_SC;
        s_src[r][c] = in[(size_t)clampi(SY0 - 4 + r, 0, H - 1) * W + clampi(SX0 - 4 + c, 0, W - 1)];
    }
    __syncthreads();
    // vertical taps of source row Y0 - 1 + r: rows -3 .. +4 around it
    for (int i = tid; i < QT_VR * QT_SC; i += 256) {
        const int r = i / QT_SC, c = i - r * QT_SC;
        int s8[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) s8[t] = s_src[r + t][c];
        s_v[r][c] = (int16_t)up_tap8(s8);
    }
    __syncthreads();
    // up sample (clamp(2Y0 + a), clamp(2X0 + b)): source sample (y, x) = that / 2.
    // A thread makes an even and an odd column of one row, so that the
    // lanes of a wave take the same parity branches (except at the picture's
    // edges, where clamping changes parities).
    auto up_val = [&](int cuy, int cux) -> int {
        const int y = cuy >> 1, x = cux >> 1;
        const int sr = y - (SY0 - 4), sc = x - (SX0 - 4), vr = y - (SY0 - 1);
        int s8[8];
        if (!(cuy & 1)) {
            if (!(cux & 1)) return s_src[sr][sc];
#pragma unroll
            for (int t = 0; t < 8; ++t) s8[t] = s_src[sr][sc - 3 + t];
            return up_tap8(s8);
        }
        if (!(cux & 1)) return s_v[vr][sc];
#pragma unroll
        for (int t = 0; t < 8; ++t) s8[t] = s_v[vr][sc - 3 + t];
        return up_tap8(s8);
    };
    // (an inner tile has SX0 == X0 and SY0 == Y0)
    const bool inner = 2 * X0 >= 0 && 2 * X0 + QT_UC <= 2 * W && 2 * Y0 >= 0 && 2 * Y0 + QT_UR <= 2 * H;
    if (inner) {
        // no up coordinate clamps: thread = (up row, segment of 8 column pairs); the
        // segment's 15 source values are read once and the 8-tap filter slides along
        constexpr int SEG = 8, NSEG = (QT_X + 1 + SEG - 1) / SEG;
        for (int i = tid; i < QT_UR * NSEG; i += 256) {
            const int a = i / NSEG, j0 = (i - a * NSEG) * SEG;
            const int16_t *row = (a & 1) ? &s_v[(a >> 1) + 1][0] : &s_src[(a >> 1) + 4][0];  // source row Y0 + a / 2
            // source column of pair j is X0 + j, stored at column j + 4 of the tile
            int16_t w[SEG + 7];
#pragma unroll
            for (int k = 0; k < SEG + 7; ++k) w[k] = row[min(j0 + 1 + k, QT_SC - 1)];
#pragma unroll
            for (int k = 0; k < SEG; ++k) {
                const int j = j0 + k;
                if (j > QT_X) break;
                s_u[a][2 * j] = w[k + 3];
                if (2 * j + 1 < QT_UC) s_u[a][2 * j + 1] = (int16_t)up_tap8_lds(&w[k]);
            }
        }
    } else {
        for (int i = tid; i < QT_UR * (QT_X + 1); i += 256) {
            const int a = i / (QT_X + 1), j = i - a * (QT_X + 1);
            const int cuy = clampi(2 * Y0 + a, 0, 2 * H - 1);
            s_u[a][2 * j] = (int16_t)up_val(cuy, clampi(2 * X0 + 2 * j, 0, 2 * W - 1));
            if (2 * j + 1 < QT_UC) s_u[a][2 * j + 1] = (int16_t)up_val(cuy, clampi(2 * X0 + 2 * j + 1, 0, 2 * W - 1));
        }
    }
    __syncthreads();
    uint32_t *__restrict__ out = reinterpret_cast<uint32_t *>(p.out);
    if (UP8) {
        // up rows 2 Y0 .. 2 Y0 + 31, columns 2 X0 .. 2 X0 + 127: dword-aligned in the plane
        const int pitch = up8_pitch(W);
        uint32_t *o = out + ((size_t)(2 * Y0 + kUp8Margin) * pitch + 2 * X0 + kUp8Margin) / 4;
        for (int i = tid; i < 2 * QT_Y * (2 * QT_X / 4); i += 256) {
            const int a = i / (2 * QT_X / 4), c = 4 * (i - a * (2 * QT_X / 4));
            o[(size_t)a * (pitch / 4) + c / 4] = (uint32_t)(s_u[a][c] + 128) | ((uint32_t)(s_u[a][c + 1] + 128) << 8) |
                                                  ((uint32_t)(s_u[a][c + 2] + 128) << 16) |
                                                  ((uint32_t)(s_u[a][c + 3] + 128) << 24);
        }
        return;
    }
    const int px = tid & (QT_X - 1);
    const int X = X0 + px;
    for (int py = tid / QT_X; py < QT_Y; py += 256 / QT_X) {
        const int Y = Y0 + py;
        if (X >= QW - 1 || Y >= QH - 1) continue;  // positions X <= W, Y <= H
        // biased samples b = u + 128; the rounded averages commute with the bias
        uint32_t b[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) b[a][c] = (uint32_t)(s_u[2 * py + a][2 * px + c] + 128);
        uint32_t *o = out + (size_t)((Y + 1) * 4) * QW + X + 1;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            // phase rows fy = 2r (one up row) and fy = 2r + 1 (rows r, r + 1)
            const uint32_t a0 = b[r][0], a1 = b[r][1], a2 = b[r][2];
            const uint32_t c0 = b[r + 1][0], c1 = b[r + 1][1], c2 = b[r + 1][2];
            const uint32_t h01 = (a0 + a1 + 1) >> 1, h12 = (a1 + a2 + 1) >> 1;
            o[(size_t)(2 * r) * QW] = a0 | (h01 << 8) | (a1 << 16) | (h12 << 24);
            const uint32_t v0 = (a0 + c0 + 1) >> 1, v1 = (a1 + c1 + 1) >> 1;
            const uint32_t q01 = (a0 + a1 + c0 + c1 + 2) >> 2, q12 = (a1 + a2 + c1 + c2 + 2) >> 2;
            o[(size_t)(2 * r + 1) * QW] = v0 | (q01 << 8) | (v1 << 16) | (q12 << 24);
        }
    }
}

// Wave copy of a clamped rectangle of an 8-bit-range plane into LDS as
// biased bytes (v + 128), rows `pitch` bytes apart; bytes [RW, padw) of each
// row are set to zero.  Lanes keep a column of a strip of at most 64.
__device__ __forceinline__ void wave_stage_u8(uint8_t *__restrict__ dst, int pitch, const int16_t *__restrict__ src,
                                              int W, int H, int x0, int y0, int RW, int RH, int padw) {
    const int lane = threadIdx.x & (WAVE - 1);
    for (int c0 = 0; c0 < padw; c0 += WAVE) {
        const int cw = min(WAVE, padw - c0);
        const int rpi = WAVE / cw;
        const int col = c0 + lane % cw, row0 = lane / cw;
        const bool on = row0 < rpi;
        const int sx = clampi(x0 + col, 0, W - 1);
        const bool inside = col < RW;
        for (int r0 = 0; r0 < RH; r0 += 16 * rpi) {
            uint8_t v[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int ry = r0 + k * rpi + row0;
                v[k] = 0;
                if (on && ry < RH && inside)
                    v[k] = (uint8_t)(src[(size_t)clampi(y0 + ry, 0, H - 1) * W + sx] + 128);
            }
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int ry = r0 + k * rpi + row0;
                if (on && ry < RH) dst[ry * pitch + col] = v[k];
            }
        }
    }
}

// SAD of an xl x yl block (packed bytes, rows of pp dwords, zero beyond xl)
// against packed byte rows of rp dwords starting at byte bo of row row0.
// MASK false: xl is a multiple of 16, no region byte needs masking.
template <bool MASK = true>
__device__ __forceinline__ int packed_sad_u8(const uint32_t *__restrict__ pic, int pp, const uint32_t *__restrict__ reg,
                                             int rp, int bo, int row0, int xl, int yl) {
    const int nd = (xl + 3) >> 2, nch = (nd + 3) >> 2;
    uint32_t m[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int valid = min(max(xl - 4 * (4 * (nch - 1) + i), 0), 4);
        m[i] = valid >= 4 ? 0xffffffffu : ((1u << (8 * valid)) - 1u);
    }
    const uint32_t *rb = reg + row0 * rp + (bo >> 2);
    const unsigned sh = bo & 3;
    uint32_t acc = 0;
    for (int y = 0; y < yl; ++y) {
        const uint32_t *pr = pic + y * pp;
        const uint32_t *rr = rb + y * rp;
        for (int c = 0; c < nch; ++c) {
            const uint32_t r0 = rr[4 * c], r1 = rr[4 * c + 1], r2 = rr[4 * c + 2], r3 = rr[4 * c + 3],
                           r4 = rr[4 * c + 4];
            uint32_t a0 = __builtin_amdgcn_alignbyte(r1, r0, sh), a1 = __builtin_amdgcn_alignbyte(r2, r1, sh);
            uint32_t a2 = __builtin_amdgcn_alignbyte(r3, r2, sh), a3 = __builtin_amdgcn_alignbyte(r4, r3, sh);
            if (MASK && c == nch - 1) {
                a0 &= m[0];
                a1 &= m[1];
                a2 &= m[2];
                a3 &= m[3];
            }
            acc = __builtin_amdgcn_sad_u8(pr[4 * c], a0, acc);
            acc = __builtin_amdgcn_sad_u8(pr[4 * c + 1], a1, acc);
            acc = __builtin_amdgcn_sad_u8(pr[4 * c + 2], a2, acc);
            acc = __builtin_amdgcn_sad_u8(pr[4 * c + 3], a3, acc);
        }
    }
    return (int)acc;
}

// packed_sad_u8 for a whole 16 x 16 block with compile-time row pitches:
// every LDS read takes an immediate offset and the rows need no loop
// control.
template <int PP, int RP>
__device__ __forceinline__ int sad16x16_u8(const uint32_t *__restrict__ pic, const uint32_t *__restrict__ reg, int bo,
                                           int row0) {
    const uint32_t *rb = reg + row0 * RP + (bo >> 2);
    const unsigned sh = bo & 3;
    uint32_t acc = 0;
#pragma unroll
    for (int y = 0; y < 16; ++y) {
        const uint32_t *pr = pic + y * PP, *rr = rb + y * RP;
        const uint32_t r0 = rr[0], r1 = rr[1], r2 = rr[2], r3 = rr[3], r4 = rr[4];
        acc = __builtin_amdgcn_sad_u8(pr[0], __builtin_amdgcn_alignbyte(r1, r0, sh), acc);
        acc = __builtin_amdgcn_sad_u8(pr[1], __builtin_amdgcn_alignbyte(r2, r1, sh), acc);
        acc = __builtin_amdgcn_sad_u8(pr[2], __builtin_amdgcn_alignbyte(r3, r2, sh), acc);
        acc = __builtin_amdgcn_sad_u8(pr[3], __builtin_amdgcn_alignbyte(r4, r3, sh), acc);
    }
    return (int)acc;
}

// ---------------------------------------------------------------------------
// Memory-level parallelism.  A wavefront step is a short chain of dependent
// operations, so its global loads must be issued together, not one per loop
// iteration: the copies below keep 32 loads in flight per lane.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void copy_ints(int *__restrict__ dst, const int *__restrict__ src, int n) {
    for (int q0 = 0; q0 < n; q0 += 32) {
        int v[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) v[i] = (q0 + i < n) ? src[q0 + i] : 0;
#pragma unroll
        for (int i = 0; i < 32; ++i)
            if (q0 + i < n) dst[q0 + i] = v[i];
    }
}

// Wave-cooperative copy of a clamped rectangle of a plane into LDS:
// dst[ry * pitch + rx] = src[clamp(y0 + ry)][clamp(x0 + rx)], 16 loads in
// flight per lane.
__device__ __forceinline__ void wave_copy_rect(int16_t *__restrict__ dst, int pitch, const int16_t *__restrict__ src,
                                               int W, int H, int x0, int y0, int RW, int RH) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int n = RW * RH;
    for (int b0 = 0; b0 < n; b0 += 16 * WAVE) {
        int16_t v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int i = b0 + k * WAVE + lane;
            v[k] = 0;
            if (i < n) {
                const int ry = i / RW, rx = i - ry * RW;
                v[k] = src[(size_t)clampi(y0 + ry, 0, H - 1) * W + clampi(x0 + rx, 0, W - 1)];
            }
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int i = b0 + k * WAVE + lane;
            if (i < n) {
                const int ry = i / RW;
                dst[ry * pitch + i - ry * RW] = v[k];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Packed SADs.  Samples are staged in LDS as v + 32768, two per dword, so that
// v_sad_u16 adds two exact absolute differences per instruction; a candidate
// row that starts at an odd sample is realigned with v_alignbyte.
// ---------------------------------------------------------------------------
// Wave copy of a clamped rectangle into LDS as biased 16-bit samples, rows
// `pitch` samples apart; samples [RW, padw) of each row are set to zero.
// Each lane keeps one column of a strip of at most 64 columns and steps over
// rows, 16 loads in flight.
__device__ __forceinline__ void wave_stage_u16(uint16_t *__restrict__ dst, int pitch, const int16_t *__restrict__ src,
                                               int W, int H, int x0, int y0, int RW, int RH, int padw) {
    const int lane = threadIdx.x & (WAVE - 1);
    for (int c0 = 0; c0 < padw; c0 += WAVE) {
        const int cw = min(WAVE, padw - c0);
        const int rpi = WAVE / cw;  // rows per pass of the wave
        const int col = c0 + lane % cw, row0 = lane / cw;
        const bool on = row0 < rpi;
        const int sx = clampi(x0 + col, 0, W - 1);
        const bool inside = col < RW;
        for (int r0 = 0; r0 < RH; r0 += 16 * rpi) {
            uint16_t v[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int ry = r0 + k * rpi + row0;
                v[k] = 0;
                if (on && ry < RH && inside)
                    v[k] = (uint16_t)(src[(size_t)clampi(y0 + ry, 0, H - 1) * W + sx] + 32768);
            }
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int ry = r0 + k * rpi + row0;
                if (on && ry < RH) dst[ry * pitch + col] = v[k];
            }
        }
    }
}

// SAD of an xl x yl block (packed rows of pp dwords, zero beyond xl) against
// the packed region rows of rp dwords starting at sample (ox, oy).  MASK
// false: xl is a multiple of 8, no region sample needs masking.
template <bool MASK = true>
__device__ __forceinline__ int packed_sad(const uint32_t *__restrict__ pic, int pp, const uint32_t *__restrict__ reg,
                                          int rp, int ox, int oy, int xl, int yl) {
    const int nd = (xl + 1) >> 1, nch = (nd + 3) >> 2;
    const int full = xl >> 1;  // dwords holding two samples of the block
    uint32_t m[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gi = 4 * (nch - 1) + i;
        m[i] = gi < full ? 0xffffffffu : (gi == full && (xl & 1) ? 0x0000ffffu : 0u);
    }
    const uint32_t *rb = reg + oy * rp + (ox >> 1);
    const unsigned sh = (ox & 1) * 2;
    uint32_t acc = 0;
    for (int y = 0; y < yl; ++y) {
        const uint32_t *pr = pic + y * pp;
        const uint32_t *rr = rb + y * rp;
        for (int c = 0; c < nch; ++c) {
            const uint32_t r0 = rr[4 * c], r1 = rr[4 * c + 1], r2 = rr[4 * c + 2], r3 = rr[4 * c + 3],
                           r4 = rr[4 * c + 4];
            uint32_t a0 = __builtin_amdgcn_alignbyte(r1, r0, sh), a1 = __builtin_amdgcn_alignbyte(r2, r1, sh);
            uint32_t a2 = __builtin_amdgcn_alignbyte(r3, r2, sh), a3 = __builtin_amdgcn_alignbyte(r4, r3, sh);
            if (MASK && c == nch - 1) {
                a0 &= m[0];
                a1 &= m[1];
                a2 &= m[2];
                a3 &= m[3];
            }
            acc = __builtin_amdgcn_sad_u16(pr[4 * c], a0, acc);
            acc = __builtin_amdgcn_sad_u16(pr[4 * c + 1], a1, acc);
            acc = __builtin_amdgcn_sad_u16(pr[4 * c + 2], a2, acc);
            acc = __builtin_amdgcn_sad_u16(pr[4 * c + 3], a3, acc);
        }
    }
    return (int)acc;
}

// ---------------------------------------------------------------------------
// Byte planes.  The level-0 planes (the pictures themselves, whose samples
// are checked to the 8-bit range at registration) are also kept as biased
// bytes v + 128 with B8_MG replicated samples on every side: rows of bp bytes
// (a multiple of 64), sample (0, 0) at byte B8_MG * bp + B8_MG of the
// allocation, every byte of a padded row holding the edge-clamped sample of
// its column.  A rectangle that stays within the margin is then copied to LDS
// as whole dwords, with no clamping.
// ---------------------------------------------------------------------------
#define B8_MG kB8Margin

// One thread per dword of the padded planes of a launch (PlaneTask: in = the
// int16 plane, out = the byte allocation, in_w = width, ow = bp, oh = height;
// tiles of 256 dwords).
__global__ __launch_bounds__(256) void k_bytes(const PlaneTask *__restrict__ pt, int npt, int total) {
    const int tile = blockIdx.x;
    if (tile >= total) return;
    int lo = 0, hi = npt - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pt[mid].offset <= tile) lo = mid;
        else hi = mid - 1;
    }
    const PlaneTask p = pt[lo];
    const int W = p.in_w, H = p.oh, dpr = p.ow >> 2;
    const int idx = (tile - p.offset) * 256 + (int)threadIdx.x;
    if (idx >= (H + 2 * B8_MG) * dpr) return;
    const int ry = idx / dpr, cx = idx - ry * dpr;
    const int16_t *__restrict__ row = p.in + (size_t)clampi(ry - B8_MG, 0, H - 1) * W;
    uint32_t d = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) d |= (uint32_t)(uint8_t)(row[clampi(cx * 4 + k - B8_MG, 0, W - 1)] + 128) << (8 * k);
    reinterpret_cast<uint32_t *>(p.out)[idx] = d;
}

// A rectangle [x0, x0 + rw) x [y0, y0 + rh) of a byte plane to be copied to
// LDS rows of `pitch` dwords: nd dwords per row, byte i of a row holding
// column x0 + i, bytes from rw on zero.  Each lane keeps one dword column of
// a band of rpp = 64 / nd rows and steps down by rpp rows per pass.
struct B8Rect {
    const uint8_t *src;  // dword-aligned byte of (x0 & ~3, y0)
    uint32_t *dst;
    int pitch, rh, sh, rpp;
    int row0, col;  // this lane's first row and its dword column
    uint32_t mask;  // bytes of the lane's dword that lie below rw
    bool on;        // lane < rpp * nd
};

__device__ __forceinline__ B8Rect b8_rect(const uint8_t *base, int bp, uint32_t *dst, int pitch, int x0, int y0, int rw,
                                          int rh, int padw) {
    const int lane = threadIdx.x & (WAVE - 1);
    B8Rect r;
    r.sh = x0 & 3;
    r.src = base + (ptrdiff_t)y0 * bp + (x0 - r.sh);
    r.dst = dst;
    r.pitch = pitch;
    r.rh = rh;
    const int nd = (padw + 3) >> 2;
    r.rpp = WAVE / nd;
    // lane / nd as a float product: exact for lane < 64, nd <= 64
    r.row0 = (int)(((float)lane + 0.5f) * (1.0f / (float)nd));
    r.col = lane - r.row0 * nd;
    r.on = r.row0 < r.rpp;
    const int valid = rw - 4 * r.col;
    r.mask = valid >= 4 ? 0xffffffffu : (valid <= 0 ? 0u : ((1u << (8 * valid)) - 1u));
    return r;
}

// True when every byte the dword copy of the rectangle reads lies in the
// padded plane: columns [-B8_MG, bp - B8_MG), rows [-B8_MG, H + B8_MG).
__device__ __forceinline__ bool b8_inside(int bp, int H, int x0, int y0, int rh, int padw) {
    return x0 >= -B8_MG && y0 >= -B8_MG && y0 + rh <= H + B8_MG && x0 + 4 * ((padw + 3) >> 2) + 4 <= bp - B8_MG;
}

// Passes p0 .. p0 + 3 of a rectangle: two aligned dword loads per element...
__device__ __forceinline__ void b8_load(const B8Rect &r, int bp, int p0, uint32_t lo[4], uint32_t hi[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int row = r.row0 + (p0 + k) * r.rpp;
        lo[k] = hi[k] = 0;
        if (r.on && row < r.rh) {
            const uint32_t *q = reinterpret_cast<const uint32_t *>(r.src + row * bp) + r.col;
            lo[k] = q[0];
            hi[k] = q[1];
        }
    }
}

// ... realigned to the rectangle's first column and masked to its width.
__device__ __forceinline__ void b8_store(const B8Rect &r, int p0, const uint32_t lo[4], const uint32_t hi[4]) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int row = r.row0 + (p0 + k) * r.rpp;
        if (r.on && row < r.rh) r.dst[row * r.pitch + r.col] = __builtin_amdgcn_alignbyte(hi[k], lo[k], r.sh) & r.mask;
    }
}

// Copy of a rectangle that leaves the margin: bytes with clamped coordinates.
__device__ __forceinline__ void b8_copy_clamped(const B8Rect &r, const uint8_t *base, int bp, int W, int H, int x0,
                                                int y0) {
    if (!r.on) return;
    for (int row = r.row0; row < r.rh; row += r.rpp) {
        const uint8_t *src = base + (ptrdiff_t)clampi(y0 + row, 0, H - 1) * bp;
        uint32_t v = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) v |= (uint32_t)src[clampi(x0 + 4 * r.col + k, 0, W - 1)] << (8 * k);
        r.dst[row * r.pitch + r.col] = v & r.mask;
    }
}

// Copies up to three rectangles of byte planes (rect i used if i < nr), the
// loads of four passes of every rectangle issued before any store.
__device__ __forceinline__ void b8_stage(const B8Rect *rc, const bool *fast, int nr, const uint8_t *const *base, int bp,
                                         int W, int H, const int *x0, const int *y0) {
    int np = 0;  // passes of the tallest fast rectangle
#pragma unroll
    for (int q = 0; q < 3; ++q)
        if (q < nr && fast[q]) np = max(np, (rc[q].rh + rc[q].rpp - 1) / rc[q].rpp);
    for (int p0 = 0; p0 < np; p0 += 4) {
        uint32_t lo[3][4], hi[3][4];
#pragma unroll
        for (int q = 0; q < 3; ++q)
            if (q < nr && fast[q]) b8_load(rc[q], bp, p0, lo[q], hi[q]);
#pragma unroll
        for (int q = 0; q < 3; ++q)
            if (q < nr && fast[q]) b8_store(rc[q], p0, lo[q], hi[q]);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q)
        if (q < nr && !fast[q]) b8_copy_clamped(rc[q], base[q], bp, W, H, x0[q], y0[q]);
}

// ---------------------------------------------------------------------------
// Pel search, bulk part.  For every block of a level, the SADs of every
// candidate within R = r + 1 (r = min(level+1, 5), pixel_match.cpp:258-259)
// of the zero vector and of the guide vector (levels below the coarsest) are
// computed in LDS.  The zero and guide candidate lists of the reference
// (pixel_match.cpp:304, 351-364) are the radius-r windows at the two centres,
// so their first minimum (in enumeration order) is taken here.  So is the
// first minimum of every r-window centred within one pel of zero or guide:
// the prediction list of most blocks is one of them, and the wavefront then
// answers it with one lookup.  One wave per block; picture block and both
// reference regions are staged in LDS.
// ---------------------------------------------------------------------------
#define PEL_MAXB 32                        // largest block side (host-checked)
#define PEL_MAXR 6
#define PEL_REG (PEL_MAXB + 2 * PEL_MAXR)  // 44
#define PEL_MAXTS ((2 * PEL_MAXR + 1) * (2 * PEL_MAXR + 1))

__device__ __forceinline__ int pel_radius(int level) { return min(level + 1, 5); }
// Pel record of a block (PEL_REC ints, blocks in wavefront order):
//   [0] zero/guide-window best SAD, [1..2] its vector, [3..4] guide vector,
//   [8 + t*9 + (dy+1)*3 + dx+1] key of the prediction window centred at
//   (dx, dy) from the centre of table t (0: zero, 1: guide), for the nine
//   predictions within one pel of the centre.  A key is SAD * 128 + k with k
//   the candidate's index in AddNewVlist order, so the smallest key is the
//   first strict-< minimum of FindBestMatchPel.
#define PEL_REC 28

__device__ __forceinline__ int wave_min_i(int v) {
    for (int m = 32; m >= 1; m >>= 1) v = min(v, __shfl_xor(v, m, WAVE));
    return v;
}

#define WAVE_SYNC()                                             \
    do {                                                        \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");  \
        __builtin_amdgcn_wave_barrier();                        \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");  \
    } while (0)

// Smallest key SAD * 128 + k over the radius-RAD window centred at tb (a
// table of row pitch 2 * RAD + 3), k the candidate's index in AddNewVlist
// order (block_match.cpp:89-126): 0 the centre, 2i-1 and 2i at (+i, 0) and
// (-i, 0), then for each row j = 1 .. RAD the pairs (x, +j), (x, -j) for x =
// -RAD .. RAD.  Keys are distinct, so the scan order does not matter.
template <int RAD>
__device__ __forceinline__ int pel_window_min(const int *tb) {
    constexpr int D = 2 * RAD + 3;
    int v = tb[0] * 128;
#pragma unroll
    for (int i = 1; i <= RAD; ++i) {
        v = min(v, tb[i] * 128 + 2 * i - 1);
        v = min(v, tb[-i] * 128 + 2 * i);
    }
#pragma unroll
    for (int j = 1; j <= RAD; ++j)
#pragma unroll
        for (int x = -RAD; x <= RAD; ++x) {
            const int k = 2 * RAD + 1 + (j - 1) * 2 * (2 * RAD + 1) + 2 * (x + RAD);
            v = min(v, tb[j * D + x] * 128 + k);
            v = min(v, tb[-j * D + x] * 128 + k + 1);
        }
    return v;
}

// LDS of one wave of k_pel_bulk, in dwords, laid out by the host for the
// largest block and radius of the launch: the picture block (rows of pp
// dwords), the two reference regions (rows of rp dwords, rp odd so that
// rows start in different banks; `keys` dwords in all), then the SAD tables.
struct PelLds {
    int pp, rp, pic, reg, keys, tab;
    __device__ int wave() const { return pic + keys + tab; }
};

// U8: the level's planes hold 8-bit-range samples (level 0: the pictures
// themselves), staged from the byte planes for v_sad_u8; otherwise (pyramid
// levels, whose 12-tap filter can overshoot that range) as 16-bit samples for
// v_sad_u16.
// The guide of block `local` of a level below the coarsest: twice the vector
// of the block above it (clamped to the coarser level's grid).
__device__ __forceinline__ dme_mv pel_guide(const JobDesc &J, int r, int level, int local) {
    const int X = J.lxb[level], xb = local % X, yb = local / X;
    const int GX = J.lxb[level + 1], GY = J.lyb[level + 1];
    dme_mv g = J.lvec[r][level + 1][(size_t)clampi(yb >> 1, 0, GY - 1) * GX + clampi(xb >> 1, 0, GX - 1)];
    g.x *= 2;
    g.y *= 2;
    return g;
}

// Blocks of k_pel_bulk per wave (1: no loop).  With more, a wave looks its
// task up once and loads the guides of all its blocks together.
#ifndef PW_PEL_BULK
#define PW_PEL_BULK 1
#endif

template <bool U8>
__device__ __forceinline__ void pel_bulk_block(const JobDesc &J, int r, int local, int level, dme_mv g,
                                               const PelLds &ly, uint32_t *__restrict__ s_pic,
                                               uint32_t *__restrict__ s_reg0, uint32_t *__restrict__ s_reg1,
                                               int *__restrict__ s_tab) {
    const int lane = threadIdx.x & 63;
    const int X = J.lxb[level];
    const int xb = local % X, yb = local / X;
    const int W = J.lw[level], H = J.lh[level];
    const Blk b = block_limits(J.blen[2], J.bsep[2], J.boff[2], W, H, xb, yb);
    const int rad = pel_radius(level), R = rad + 1, D = 2 * R + 1, TS = D * D;
    const bool has_guide = level < J.depth;
    const int ntab = has_guide ? 2 : 1;
    if (!has_guide) g.x = g.y = 0;
    int *__restrict__ rec = J.ptab[r] + (size_t)wf_pos(xb, yb, X, J.lyb[level]) * PEL_REC;
    int best_sad = 10000000;
    dme_mv best_mv;
    best_mv.x = 0;
    best_mv.y = 0;
    int mykey = INT_MAX;  // lane t*9 + w keeps the key of window w of table t
    if (!b.empty()) {
        const int RW = b.xl + 2 * R, RH = b.yl + 2 * R;
        if (U8) {
            const int bp = J.p8;
            const uint8_t *const base[3] = {J.pic8, J.ref8[r], J.ref8[r]};
            const int x0[3] = {b.xp, b.xp - R, b.xp + g.x - R}, y0[3] = {b.yp, b.yp - R, b.yp + g.y - R};
            const int padw[3] = {(b.xl + 15) & ~15, RW, RW};
            B8Rect rc[3];
            bool fast[3];
            rc[0] = b8_rect(base[0], bp, s_pic, ly.pp, x0[0], y0[0], b.xl, b.yl, padw[0]);
            rc[1] = b8_rect(base[1], bp, s_reg0, ly.rp, x0[1], y0[1], RW, RH, padw[1]);
            rc[2] = b8_rect(base[2], bp, s_reg1, ly.rp, x0[2], y0[2], RW, RH, padw[2]);
            fast[0] = b8_inside(bp, H, x0[0], y0[0], b.yl, padw[0]);
            fast[1] = b8_inside(bp, H, x0[1], y0[1], RH, padw[1]);
            fast[2] = b8_inside(bp, H, x0[2], y0[2], RH, padw[2]);
            b8_stage(rc, fast, has_guide ? 3 : 2, base, bp, W, H, x0, y0);
        } else {
            const int16_t *__restrict__ pic = J.pic[level];
            const int16_t *__restrict__ ref = J.ref[r][level];
            wave_stage_u16(reinterpret_cast<uint16_t *>(s_pic), 2 * ly.pp, pic, W, H, b.xp, b.yp, b.xl, b.yl,
                           (b.xl + 7) & ~7);
            wave_stage_u16(reinterpret_cast<uint16_t *>(s_reg0), 2 * ly.rp, ref, W, H, b.xp - R, b.yp - R, RW, RH, RW);
            if (has_guide)
                wave_stage_u16(reinterpret_cast<uint16_t *>(s_reg1), 2 * ly.rp, ref, W, H, b.xp + g.x - R,
                               b.yp + g.y - R, RW, RH, RW);
        }
        WAVE_SYNC();
        const bool whole = (b.xl & (U8 ? 15 : 7)) == 0;  // no partial dword chunk
        // the common level-0 case (16 x 16 blocks, radius 1) with fixed pitches
        const bool b16 = U8 && b.xl == 16 && b.yl == 16 && ly.pp == 4 && ly.rp == 7;
        for (int c = lane; c < ntab * TS; c += WAVE) {
            const int t = c >= TS ? 1 : 0, e = c - t * TS;
            const int oy = e / D, ox = e - oy * D;  // region offset of the candidate
            const uint32_t *reg = t ? s_reg1 : s_reg0;
            if (U8)
                s_tab[c] = b16 ? sad16x16_u8<4, 7>(s_pic, reg, ox, oy)
                               : (whole ? packed_sad_u8<false>(s_pic, ly.pp, reg, ly.rp, ox, oy, b.xl, b.yl)
                                        : packed_sad_u8<true>(s_pic, ly.pp, reg, ly.rp, ox, oy, b.xl, b.yl));
            else
                s_tab[c] = whole ? packed_sad<false>(s_pic, ly.pp, reg, ly.rp, ox, oy, b.xl, b.yl)
                                 : packed_sad<true>(s_pic, ly.pp, reg, ly.rp, ox, oy, b.xl, b.yl);
        }
        WAVE_SYNC();
        // first minima (key = SAD * 128 + enumeration index) of the nine
        // windows around each centre: lane t*9 + w scans window w of table t
        if (lane < ntab * 9) {
            const int t = lane >= 9 ? 1 : 0, w = lane - 9 * t;
            const int *tb = s_tab + t * TS + (w / 3 - 1 + R) * D + (w % 3 - 1 + R);
            switch (rad) {
                case 1: mykey = pel_window_min<1>(tb); break;
                case 2: mykey = pel_window_min<2>(tb); break;
                case 3: mykey = pel_window_min<3>(tb); break;
                case 4: mykey = pel_window_min<4>(tb); break;
                default: mykey = pel_window_min<5>(tb); break;
            }
        }
        // zero list, then guide list (radius rad, centred windows)
        const int k0 = __shfl(mykey, 4, WAVE), k1 = __shfl(mykey, 13, WAVE);
        best_sad = k0 >> 7;
        best_mv = win_off(k0 & 127, rad);
        if (has_guide && (k1 >> 7) < best_sad) {
            best_sad = k1 >> 7;
            best_mv = win_off(k1 & 127, rad);
            best_mv.x += g.x;
            best_mv.y += g.y;
        }
    }
    if (lane < 18) rec[8 + lane] = mykey;
    if (lane == 0) {
        rec[0] = best_sad;
        rec[1] = best_mv.x;
        rec[2] = best_mv.y;
        rec[3] = g.x;
        rec[4] = g.y;
    }
}

template <bool U8>
__global__ __launch_bounds__(256) void k_pel_bulk(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                  int ntasks, int total, int level, PelLds ly) {
    extern __shared__ int s_dyn[];
    const int wv = wave_id(), lane = threadIdx.x & 63;
    uint32_t *__restrict__ s_pic = reinterpret_cast<uint32_t *>(s_dyn) + wv * ly.wave();
    uint32_t *__restrict__ s_reg0 = s_pic + ly.pic;
    uint32_t *__restrict__ s_reg1 = s_reg0 + ly.reg;
    int *__restrict__ s_tab = reinterpret_cast<int *>(s_pic + ly.pic + ly.keys);
#if PW_PEL_BULK > 1
    const int g0 = (xcd_block() * 4 + wv) * PW_PEL_BULK;
    if (g0 >= total) return;  // each wave works alone
    const int g1 = min(g0 + PW_PEL_BULK, total);
    int ti = find_task_wave(tasks, ntasks, g0);
    // lane k < g1 - g0: the guide of block g0 + k (its own task lookup)
    int gx = 0, gy = 0;
    if (lane < g1 - g0) {
        const int gi = g0 + lane;
        int t = ti;
        while (t + 1 < ntasks && tasks[t + 1].offset <= gi) ++t;
        const Task tk = tasks[t];
        const JobDesc &J = jobs[tk.job];
        if (level < J.depth) {
            const dme_mv g = pel_guide(J, tk.ref, level, gi - tk.offset);
            gx = g.x;
            gy = g.y;
        }
    }
#pragma unroll 1
    for (int gi = g0; gi < g1; ++gi) {
        while (ti + 1 < ntasks && tasks[ti + 1].offset <= gi) ++ti;
        const Task tk = tasks[ti];
        dme_mv g;
        g.x = __builtin_amdgcn_readlane(gx, gi - g0);
        g.y = __builtin_amdgcn_readlane(gy, gi - g0);
        WAVE_SYNC();  // the previous block no longer reads the wave's LDS
        pel_bulk_block<U8>(jobs[tk.job], tk.ref, gi - tk.offset, level, g, ly, s_pic, s_reg0, s_reg1, s_tab);
    }
#else
    const int gi = xcd_block() * 4 + wv;
    if (gi >= total) return;  // each wave works alone
    const Task tk = tasks[find_task_wave(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    dme_mv g;
    g.x = g.y = 0;
    if (level < J.depth) g = pel_guide(J, tk.ref, level, gi - tk.offset);
    pel_bulk_block<U8>(J, tk.ref, gi - tk.offset, level, g, ly, s_pic, s_reg0, s_reg1, s_tab);
#endif
}

// ---------------------------------------------------------------------------
// Full search, table part (PixelMatcher::MatchPic with FullSearch,
// pixel_match.cpp:91-155, 249-264, 365-370): no pyramid; at level 0 the zero
// list is the whole window of radius (xr, yr) = min(tdiff, 3) * (x, y range)
// and the prediction list has radius 1.  One workgroup per (picture,
// reference, block) stages the block and its reference window once, as
// biased bytes, and every thread takes candidates of the window; the first
// strict-< minimum in AddNewVlist order is the smallest (SAD, index) pair.
// The record is the level-0 record of k_pel_bulk (no guide): the zero-list
// winner and the keys of the nine radius-1 windows centred within one pel of
// zero, which k_pel_wave consumes unchanged.
// ---------------------------------------------------------------------------
// index of (dx, dy) in AddNewVlist's order over the (2xr+1) x (2yr+1) window
__device__ __forceinline__ int full_enum(int dx, int dy, int xr) {
    if (dy == 0) return dx == 0 ? 0 : (dx > 0 ? 2 * dx - 1 : -2 * dx);
    return 1 + 2 * xr + (abs(dy) - 1) * 2 * (2 * xr + 1) + 2 * (dx + xr) + (dy < 0 ? 1 : 0);
}

// Whole blocks of NW dwords x ROWS rows (the common case): a thread takes
// groups of four horizontally adjacent candidates, which share each
// reference row's NW + 2 dword loads (candidate k of a group is the row
// realigned by k more bytes) and each block row's loads (one LDS broadcast).
// Returns the smallest (SAD, index) key of the thread's candidates.
template <int NW, int ROWS>
__device__ __forceinline__ unsigned long long full_groups(const uint32_t *__restrict__ s_pic,
                                                          const uint32_t *__restrict__ s_reg, int rp, int rx, int ry,
                                                          int xr, int yr, int tid) {
    constexpr int PP = ((NW + 3) / 4) * 4;  // block row pitch in LDS (chunks of 16 bytes)
    const int WX = 2 * xr + 1, ngx = (WX + 3) >> 2, ng = ngx * (2 * yr + 1);
    unsigned long long best = ~0ull;
    for (int gidx = tid; gidx < ng; gidx += 256) {
        const int gy = gidx / ngx, dy = gy - yr, dx0 = (gidx - gy * ngx) * 4 - xr;
        const int bo = dx0 + rx;  // byte offset of the group's first candidate in its region row
        const uint32_t *__restrict__ rb = s_reg + (dy + ry) * rp + (bo >> 2);
        const unsigned s0 = bo & 3;
        uint32_t acc[4] = {0, 0, 0, 0};
#pragma unroll 4
        for (int y = 0; y < ROWS; ++y) {
            uint32_t r[NW + 2];
#pragma unroll
            for (int j = 0; j < NW + 2; ++j) r[j] = rb[y * rp + j];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const unsigned t = s0 + k, o = t >> 2, sh = t & 3;
#pragma unroll
                for (int j = 0; j < NW; ++j) {
                    const uint32_t lo = o ? r[j + 1] : r[j], hi = o ? r[j + 2] : r[j + 1];
                    acc[k] = __builtin_amdgcn_sad_u8(s_pic[y * PP + j], __builtin_amdgcn_alignbyte(hi, lo, sh), acc[k]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int dx = dx0 + k;
            if (dx > xr) break;
            const unsigned long long key = ((unsigned long long)acc[k] << 32) | (unsigned)full_enum(dx, dy, xr);
            best = key < best ? key : best;
        }
    }
    return best;
}

// LDS of k_pel_full, in dwords: per-wave minima, the 5 x 5 SAD table around
// zero, the block (rows of pp), the window (rows of rp, with margins of
// rx, ry >= 2 so that the 5 x 5 neighbourhood of zero is staged too).
struct FullLds {
    int pp, rp, rows, pic;
};

__global__ __launch_bounds__(256) void k_pel_full(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                  int ntasks, int total, FullLds ly) {
    extern __shared__ int s_dyn[];
    unsigned long long *__restrict__ s_min = reinterpret_cast<unsigned long long *>(s_dyn);  // 4, 8-byte aligned
    int *__restrict__ s_t25 = s_dyn + 8;
    uint32_t *__restrict__ s_pic = reinterpret_cast<uint32_t *>(s_t25 + 26);
    uint32_t *__restrict__ s_reg = s_pic + ly.pic;
    const int gi = blockIdx.x;
    if (gi >= total) return;
    const Task tk = tasks[find_task_wave(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    const int r = tk.ref, local = gi - tk.offset;
    const int X = J.lxb[0], Y = J.lyb[0];
    const int xb = local % X, yb = local / X;
    const int W = J.lw[0], H = J.lh[0];
    const Blk b = block_limits(J.blen[2], J.bsep[2], J.boff[2], W, H, xb, yb);
    int *__restrict__ rec = J.ptab[r] + (size_t)wf_pos(xb, yb, X, Y) * PEL_REC;
    const int tid = threadIdx.x, wv = wave_id(), lane = tid & 63;
    if (b.empty()) {
        if (tid < PEL_REC) rec[tid] = tid == 0 ? 10000000 : (tid >= 8 && tid < 26 ? INT_MAX : 0);
        return;
    }
    const int xr = J.fxr[r], yr = J.fyr[r];
    const int rx = max(xr, 2), ry = max(yr, 2);
    const int RW = b.xl + 2 * rx, RH = b.yl + 2 * ry;
    const int pp = ((b.xl + 15) >> 4) * 4, rp = ly.rp;
    // stage: wave 0 the block, every wave a quarter of the window's rows
    if (wv == 0) wave_stage_u8(reinterpret_cast<uint8_t *>(s_pic), 4 * pp, J.pic[0], W, H, b.xp, b.yp, b.xl, b.yl, 4 * pp);
    {
        const int q = (RH + 3) >> 2, r0 = wv * q, nr = min(q, RH - r0);
        if (nr > 0)
            wave_stage_u8(reinterpret_cast<uint8_t *>(s_reg + r0 * rp), 4 * rp, J.ref[r][0], W, H, b.xp - rx,
                          b.yp - ry + r0, RW, nr, 4 * rp);
    }
    __syncthreads();
    const bool whole = (b.xl & 15) == 0;
    auto sad_at = [&](int dx, int dy) {
        return whole ? packed_sad_u8<false>(s_pic, pp, s_reg, rp, dx + rx, dy + ry, b.xl, b.yl)
                     : packed_sad_u8<true>(s_pic, pp, s_reg, rp, dx + rx, dy + ry, b.xl, b.yl);
    };
    const int WX = 2 * xr + 1, nc = WX * (2 * yr + 1);
    unsigned long long best = ~0ull;
    if (b.xl == 16 && b.yl == 16)
        best = full_groups<4, 16>(s_pic, s_reg, rp, rx, ry, xr, yr, tid);
    else if (b.xl == 12 && b.yl == 12)
        best = full_groups<3, 12>(s_pic, s_reg, rp, rx, ry, xr, yr, tid);
    else
        for (int c = tid; c < nc; c += 256) {
            const int dy = c / WX - yr, dx = c - (dy + yr) * WX - xr;
            const unsigned long long key =
                ((unsigned long long)(unsigned)sad_at(dx, dy) << 32) | (unsigned)full_enum(dx, dy, xr);
            best = key < best ? key : best;
        }
    if (tid < 25) s_t25[tid] = sad_at(tid % 5 - 2, tid / 5 - 2);
    for (int m = 32; m >= 1; m >>= 1) {
        const unsigned long long o = __shfl_xor(best, m, WAVE);
        best = o < best ? o : best;
    }
    if (lane == 0) s_min[wv] = best;
    __syncthreads();
    if (wv != 0) return;
    best = s_min[0];
#pragma unroll
    for (int w = 1; w < 4; ++w) best = s_min[w] < best ? s_min[w] : best;
    // keys of the radius-1 windows centred within one pel of zero
    int key = INT_MAX;
    if (lane < 9) key = pel_window_min<1>(s_t25 + (lane / 3 + 1) * 5 + (lane % 3 + 1));
    if (lane < 18) rec[8 + lane] = key;
    if (lane == 0) {
        const int k = (int)(best & 0xffffffffu);
        int dx, dy;
        if (k <= 2 * xr) {  // the centre row
            dy = 0;
            dx = k == 0 ? 0 : ((k & 1) ? (k + 1) >> 1 : -(k >> 1));
        } else {
            const int k2 = k - 1 - 2 * xr, row = k2 / (2 * WX), rem = k2 - row * 2 * WX;
            dx = (rem >> 1) - xr;
            dy = (rem & 1) ? -(row + 1) : row + 1;
        }
        rec[0] = (int)(best >> 32);
        rec[1] = dx;
        rec[2] = dy;
        rec[3] = 0;
        rec[4] = 0;
    }
}

// ---------------------------------------------------------------------------
// Pel search, wavefront part (PixelMatcher::MatchPic/DoBlock,
// pixel_match.cpp:242-380; BlockMatcher::FindBestMatchPel,
// block_match.cpp:242-275).  One workgroup per (picture, reference), one
// thread per block row: at step t row y works on block x = t - 2y, reading
// the left, top and top-right results from a 3-slot LDS ring per row.  While
// the step runs, the workgroup copies the records of diagonal t + 1 into the
// other LDS bank.  A prediction within one pel of zero or of the guide is
// answered by one lookup of its window key.  The other predictions of the
// step ("misses") are queued in LDS and the whole workgroup computes their
// windows together: it stages each block and its reference window, sums row
// SADs per candidate with LDS atomics and takes the smallest key.
// ---------------------------------------------------------------------------
struct PelMiss {
    int row, xp, yp, xl, yl, px, py, key;
};

__global__ __launch_bounds__(1024) void k_pel_wave(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                   int level, int max_rows, int miss_cap, int stage_elems) {
    extern __shared__ int s_dyn[];
    __shared__ int s_nmiss;
    const Task tk = tasks[blockIdx.x];
    const JobDesc &J = jobs[tk.job];
    if (level > J.depth || tk.ref >= J.nrefs) return;  // uniform per workgroup
    const int r = tk.ref;
    const int X = J.lxb[level], Y = J.lyb[level];
    const int W = J.lw[level], H = J.lh[level];
    const int rad = pel_radius(level);
    const int n = (2 * rad + 1) * (2 * rad + 1);
    const bool has_guide = level < J.depth;
    const int16_t *__restrict__ pic = J.pic[level];
    const int16_t *__restrict__ ref = J.ref[r][level];
    const int *__restrict__ tab = J.ptab[r];
    dme_mv *__restrict__ V = J.lvec[r][level];
    dme_cost *__restrict__ C = J.lpc[r][level];
    int blen[2] = {J.blen[2][0], J.blen[2][1]}, bsep[2] = {J.bsep[2][0], J.bsep[2][1]};
    int boff[2] = {J.boff[2][0], J.boff[2][1]};
    const int YL = blen[1];
    const int tid = threadIdx.x, nth = blockDim.x;
    // LDS: ring [max_rows][3] mv | slots [2][max_rows][PEL_REC] | queue [miss_cap] |
    //      sums [miss_cap][n] | staging [miss_cap][stage_elems] int16
    dme_mv *ring = reinterpret_cast<dme_mv *>(s_dyn);
    int *slots = s_dyn + 6 * max_rows;
    PelMiss *queue = reinterpret_cast<PelMiss *>(slots + 2 * max_rows * PEL_REC);
    int *msum = reinterpret_cast<int *>(queue + miss_cap);
    int16_t *mstage = reinterpret_cast<int16_t *>(msum + miss_cap * n);
    if (tid == 0) s_nmiss = 0;
    const int T = (X - 1) + 2 * (Y - 1);
    for (int t = -1; t <= T; ++t) {
        WfCopy<4> cp;
        wf_issue(cp, tab, slots, PEL_REC, max_rows, t + 1, X, Y);
        for (int base = 0; base < Y; base += nth) {
            const int yb = base + tid;
            const int xb = t - 2 * yb;
            const bool active = yb < Y && xb >= 0 && xb < X;
            dme_mv pred, best;
            pred.x = pred.y = 0;
            best = pred;
            int best_sad = 10000000, slot = -1;
            Blk b;
            b.xp = b.yp = 0;
            b.xl = b.yl = 0;
            __syncthreads();  // the queue counter is reset; ring and bank writes are visible
            if (active) {
                const int *sl = slots + ((xb & 1) * max_rows + yb) * PEL_REC;
                // prediction rules (pixel_match.cpp:314-347)
                if (yb == 0) {
                    if (xb > 0) pred = ring[yb * 3 + (t - 1) % 3];
                } else if (xb == 0) {
                    pred = ring[(yb - 1) * 3 + (t - 2) % 3];
                } else if (xb < X - 1) {
                    pred = med3(ring[yb * 3 + (t - 1) % 3], ring[(yb - 1) * 3 + (t - 2) % 3],
                                ring[(yb - 1) * 3 + (t - 1) % 3]);
                } else {
                    pred = mean2(ring[(yb - 1) * 3 + (t - 2) % 3], ring[yb * 3 + (t - 1) % 3]);
                }
                best_sad = sl[0];
                best.x = sl[1];
                best.y = sl[2];
                b = block_limits(blen, bsep, boff, W, H, xb, yb);
                if (!b.empty()) {
                    const int gx = sl[3], gy = sl[4];
                    int key = -1;
                    if (J.full && abs(pred.x) < J.fxr[r] && abs(pred.y) < J.fyr[r])
                        // full search: every candidate of the prediction list lies in
                        // the zero window, whose first minimum is already the best,
                        // so none is strictly smaller
                        key = INT_MAX;
                    else if (has_guide && abs(pred.x - gx) <= 1 && abs(pred.y - gy) <= 1)
                        key = sl[17 + (pred.y - gy + 1) * 3 + (pred.x - gx + 1)];
                    else if (abs(pred.x) <= 1 && abs(pred.y) <= 1)
                        key = sl[8 + (pred.y + 1) * 3 + (pred.x + 1)];
                    if (key >= 0) {
                        if ((key >> 7) < best_sad) {
                            best_sad = key >> 7;
                            best = win_off(key & 127, rad);
                            best.x += pred.x;
                            best.y += pred.y;
                        }
                    } else {
                        slot = atomicAdd(&s_nmiss, 1);
                    }
                }
            }
            __syncthreads();
            const int nmiss = s_nmiss;  // reset after the rounds; the next pass starts with a barrier
            for (int m0 = 0; m0 < nmiss; m0 += miss_cap) {
                const int nm = min(miss_cap, nmiss - m0);
                if (slot >= m0 && slot < m0 + nm) {
                    PelMiss e;
                    e.row = yb;
                    e.xp = b.xp;
                    e.yp = b.yp;
                    e.xl = b.xl;
                    e.yl = b.yl;
                    e.px = pred.x;
                    e.py = pred.y;
                    e.key = INT_MAX;
                    queue[slot - m0] = e;
                }
                __syncthreads();
                // stage each block (pitch xl) and its window (pitch xl + 2 rad)
                for (int i0 = 0; i0 < nm * stage_elems; i0 += 8 * nth) {
                    int16_t v[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const int i = i0 + k * nth + tid;
                        v[k] = 0;
                        if (i < nm * stage_elems) {
                            const int e = i / stage_elems, j = i - e * stage_elems;
                            const PelMiss q = queue[e];
                            const int nb = q.xl * q.yl, RW = q.xl + 2 * rad, RH = q.yl + 2 * rad;
                            if (j < nb) {
                                const int dy = j / q.xl;
                                v[k] = pic[(size_t)(q.yp + dy) * W + q.xp + j - dy * q.xl];
                            } else if (j - nb < RW * RH) {
                                const int j2 = j - nb, ry = j2 / RW;
                                v[k] = ref[(size_t)clampi(q.yp + q.py - rad + ry, 0, H - 1) * W +
                                           clampi(q.xp + q.px - rad + j2 - ry * RW, 0, W - 1)];
                            }
                        }
                    }
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const int i = i0 + k * nth + tid;
                        if (i < nm * stage_elems) mstage[i] = v[k];
                    }
                }
                for (int i = tid; i < nm * n; i += nth) msum[i] = 0;
                __syncthreads();
                // row SADs of every (miss, candidate, row)
                for (int i = tid; i < nm * n * YL; i += nth) {
                    const int e = i / (n * YL), rem = i - e * n * YL;
                    const int c = rem / YL, y = rem - c * YL;
                    const PelMiss q = queue[e];
                    if (y < q.yl) {
                        const dme_mv o = win_off(c, rad);
                        const int RW = q.xl + 2 * rad;
                        const int16_t *bp = mstage + e * stage_elems + y * q.xl;
                        const int16_t *rp = mstage + e * stage_elems + q.xl * q.yl + (y + o.y + rad) * RW + o.x + rad;
                        int sd = 0;
                        for (int x = 0; x < q.xl; ++x) sd += abs((int)bp[x] - (int)rp[x]);
                        atomicAdd(&msum[e * n + c], sd);
                    }
                }
                __syncthreads();
                for (int i = tid; i < nm * n; i += nth) {
                    const int e = i / n, c = i - e * n;
                    atomicMin(&queue[e].key, msum[i] * 128 + c);
                }
                __syncthreads();
                if (slot >= m0 && slot < m0 + nm) {
                    const int key = queue[slot - m0].key;
                    if ((key >> 7) < best_sad) {
                        best_sad = key >> 7;
                        best = win_off(key & 127, rad);
                        best.x += pred.x;
                        best.y += pred.y;
                    }
                }
                __syncthreads();
            }
            if (tid == 0 && nmiss) s_nmiss = 0;
            if (active) {
                const size_t k = (size_t)yb * X + xb;
                V[k] = best;
                dme_cost c;
                c.SAD = (float)best_sad;
                c.mvcost = (float)(int16_t)norm1(best.x - pred.x, best.y - pred.y);
                c.total = c.SAD + 0.0f * c.mvcost;  // SetTotal(0.0)
                C[k] = c;
                ring[yb * 3 + t % 3] = best;
            }
        }
        wf_finish(cp);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Up-sampled SADs (BlockDiffHalfPel / QuarterPel, me_utils.cpp:410-867;
// MMX simple_block_diff_up_mmx_4, me_utils_mmx.cpp:201-533).
// step 1: mv in half pel, sample up[2y+mv.y][2x+mv.x]; step 2: quarter pel,
// rounded average of 1, 2 or 4 half-pel samples; step 3: eighth pel
// (BlockDiffEighthPel, me_utils.cpp:890-1170), start mv >> 2 in the up
// plane, weights (4-rx)(4-ry), rx(4-ry), (4-rx)ry, rx ry of the remainders,
// +8 >> 4.  Clamped reads equal both reference paths' values.
// ---------------------------------------------------------------------------
struct UpStart {
    int rsx, rsy, rx, ry;
    bool bounds;
    bool e;  // eighth pel: rx, ry in 0..3
};

// E8 false: step is 1 or 2 (the forms below eighth pel compile without it).
template <bool E8 = true>
__device__ __forceinline__ UpStart up_start(const Blk &b, dme_mv mv, int step, int UW, int UH) {
    UpStart s;
    s.e = E8 && step == 3;
    if (step == 1) {
        s.rsx = (b.xp << 1) + mv.x; s.rsy = (b.yp << 1) + mv.y; s.rx = 0; s.ry = 0;
    } else if (!E8 || step == 2) {
        s.rsx = (b.xp << 1) + (mv.x >> 1); s.rsy = (b.yp << 1) + (mv.y >> 1);
        s.rx = mv.x & 1; s.ry = mv.y & 1;
    } else {
        s.rsx = (b.xp << 1) + (mv.x >> 2); s.rsy = (b.yp << 1) + (mv.y >> 2);
        s.rx = mv.x & 3; s.ry = mv.y & 3;
    }
    s.bounds = s.rsx < 0 || s.rsx + (b.xl << 1) >= UW || s.rsy < 0 || s.rsy + (b.yl << 1) >= UH;
    return s;
}

// Up-plane sample (ux, uy), clamped, from the quarter-pel planes: their even
// phases hold the up samples themselves (position (X, Y) phase (2 fx', 2 fy')
// is up sample (2X + fx', 2Y + fy')).  Biased (v + 128).
__device__ __forceinline__ int qp_raw(const uint32_t *__restrict__ qp, int W, int H, int ux, int uy) {
    ux = clampi(ux, 0, 2 * W - 1);
    uy = clampi(uy, 0, 2 * H - 1);
    const uint32_t d = qp[(size_t)(((uy >> 1) + 1) * 4 + 2 * (uy & 1)) * (W + 2) + (ux >> 1) + 1];
    return (int)((d >> (16 * (ux & 1))) & 0xffu);
}

// Eighth-pel sample (biased) at up position (ux + rx / 4, uy + ry / 4); the
// bias commutes with the weights, which sum to 16.
__device__ __forceinline__ int qp_eighth(const uint32_t *__restrict__ qp, int W, int H, int ux, int uy, int rx, int ry) {
    const int a = qp_raw(qp, W, H, ux, uy), b = qp_raw(qp, W, H, ux + 1, uy);
    const int c = qp_raw(qp, W, H, ux, uy + 1), d = qp_raw(qp, W, H, ux + 1, uy + 1);
    return ((4 - rx) * (4 - ry) * a + rx * (4 - ry) * b + (4 - rx) * ry * c + rx * ry * d + 8) >> 4;
}

// Up-plane sample of pixel (dy, dx) for start s, read from the quarter-pel
// planes: quarter-pel position (2 rsx + rx + 4 dx, 2 rsy + ry + 4 dy), or
// the eighth-pel sample at up position (rsx + 2 dx, rsy + 2 dy).
// The same sample of a combined-component job, from the int16 up plane
// (2W x 2H) itself: the weights of the four clamped up samples around it
// ((2 - rx)(2 - ry) ... + 2 >> 2; eighth pel (4 - rx)(4 - ry) ... + 8 >> 4),
// as BlockDiff{Half,Quarter,Eighth}Pel form it (me_utils.cpp:410-1170).
template <bool E8>
__device__ __forceinline__ int up16_at(const int16_t *__restrict__ up, int W, int H, const UpStart &s, int dy, int dx) {
    const int UW = 2 * W, UH = 2 * H;
    const int uy = s.rsy + 2 * dy, ux = s.rsx + 2 * dx;
    const int y0 = clampi(uy, 0, UH - 1), y1 = clampi(uy + 1, 0, UH - 1);
    const int x0 = clampi(ux, 0, UW - 1), x1 = clampi(ux + 1, 0, UW - 1);
    const bool e = E8 && s.e;
    const int one = e ? 4 : 2;
    const int w0 = (one - s.rx) * (one - s.ry), w1 = s.rx * (one - s.ry), w2 = (one - s.rx) * s.ry, w3 = s.rx * s.ry;
    const int v = w0 * up[(size_t)y0 * UW + x0] + w1 * up[(size_t)y0 * UW + x1] + w2 * up[(size_t)y1 * UW + x0] +
                  w3 * up[(size_t)y1 * UW + x1];
    return e ? (v + 8) >> 4 : (v + 2) >> 2;
}

// COMB: qp is a combined-component job's int16 up plane (up16_at).
template <bool E8, bool COMB = false>
__device__ __forceinline__ int qp_at(const uint32_t *__restrict__ qp, int W, int H, const UpStart &s, int dy, int dx) {
    if (COMB) return up16_at<E8>(reinterpret_cast<const int16_t *>(qp), W, H, s, dy, dx);
    if (E8 && s.e) return qp_eighth(qp, W, H, s.rsx + 2 * dx, s.rsy + 2 * dy, s.rx, s.ry) - 128;
    const int Xq = 2 * s.rsx + s.rx + 4 * dx, Yq = 2 * s.rsy + s.ry + 4 * dy;
    const int X = clampi(Xq >> 2, -1, W), Y = clampi(Yq >> 2, -1, H);
    const uint32_t d = qp[(size_t)((Y + 1) * 4 + (Yq & 3)) * (W + 2) + X + 1];
    return (int)((d >> (8 * (Xq & 3))) & 0xffu) - 128;
}

// Integer SADs of up to 4 candidates against the up-converted reference;
// wave-uniform.  (candidate, pixel) items are spread over the lanes, 8 per
// lane at a time with all their loads issued together.
template <bool E8, bool COMB = false>
__device__ __noinline__ void wave_up_sads(const int16_t *__restrict__ pic, int W, const uint32_t *__restrict__ qp, int UW,
                                          int UH, const Blk &b, int step, const dme_mv *cands, int nc, int *out) {
    const int lane = threadIdx.x & (WAVE - 1);
    UpStart st[4];
    for (int c = 0; c < 4; ++c) st[c] = up_start<E8>(b, cands[c < nc ? c : 0], step, UW, UH);
    const int np = b.xl * b.yl, n = nc * np;
    int acc0 = 0, acc1 = 0, acc2 = 0, acc3 = 0;
    for (int b0 = 0; b0 < n; b0 += 8 * WAVE) {
        int pv[8], rv[8], cid[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int it = b0 + k * WAVE + lane;
            cid[k] = -1;
            pv[k] = rv[k] = 0;
            if (it < n) {
                const int c = it / np, p = it - c * np;
                const int dy = p / b.xl, dx = p - dy * b.xl;
                UpStart s0 = st[0];
#pragma unroll
                for (int e = 1; e < 4; ++e)
                    if (c == e) s0 = st[e];
                cid[k] = c;
                pv[k] = pic[(size_t)(b.yp + dy) * W + b.xp + dx];
                rv[k] = qp_at<E8, COMB>(qp, UW >> 1, UH >> 1, s0, dy, dx);
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int d = abs(rv[k] - pv[k]);
            acc0 += cid[k] == 0 ? d : 0;
            acc1 += cid[k] == 1 ? d : 0;
            acc2 += cid[k] == 2 ? d : 0;
            acc3 += cid[k] == 3 ? d : 0;
        }
    }
    if (nc > 0) out[0] = wave_sum(acc0);
    if (nc > 1) out[1] = wave_sum(acc1);
    if (nc > 2) out[2] = wave_sum(acc2);
    if (nc > 3) out[3] = wave_sum(acc3);
}

// Per-pixel float accumulation of the bounds-checked path (and of the plain
// C build), raster order (me_utils.cpp:532-541, 843-867), computed by the
// whole wave (uniform arguments).  For
// 0 <= s < 2^24 and an integer d >= 0, s + d is exact while it stays below
// the power of two above s: both are multiples of the spacing of s's binade.
// So the chain only rounds where a partial sum enters a higher binade.  The
// wave takes 64 differences at a time, forms their prefix sums, adds the runs
// between crossings in one exact step each and performs the crossing
// additions one by one.  Sums from 2^24 up are added one pixel at a time.
__device__ __forceinline__ float binade_limit(float s) {
    const unsigned e = __float_as_uint(s) >> 23;  // s >= 0
    return __uint_as_float((e + 1) << 23);
}

template <bool E8, bool COMB = false>
__device__ __noinline__ float wave_chain(float start, const int16_t *__restrict__ pic, int W,
                                        const uint32_t *__restrict__ qp, int UW, int UH, const Blk &b, const UpStart &st) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int n = b.xl * b.yl;
    float s = start;
    for (int b0 = 0; b0 < n; b0 += 4 * WAVE) {
        int d[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = b0 + k * WAVE + lane;
            const int ii = i < n ? i : n - 1;
            const int dy = ii / b.xl, dx = ii - dy * b.xl;
            const int v =
                abs(qp_at<E8, COMB>(qp, UW >> 1, UH >> 1, st, dy, dx) - (int)pic[(size_t)(b.yp + dy) * W + b.xp + dx]);
            d[k] = i < n ? v : 0;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (b0 + k * WAVE >= n) break;
            int x = d[k];
#pragma unroll
            for (int off = 1; off < WAVE; off <<= 1) {
                const int y = __shfl_up(x, off, WAVE);
                if (lane >= off) x += y;
            }
            const int total = __builtin_amdgcn_readlane(x, WAVE - 1);
            int consumed = 0, first = 0;
            while (first < WAVE) {
                if (s >= 16777216.0f) {
                    for (int j = first; j < WAVE; ++j)
                        s += (float)(__builtin_amdgcn_readlane(x, j) - (j ? __builtin_amdgcn_readlane(x, j - 1) : 0));
                    break;
                }
                const double D = (double)binade_limit(s) - (double)s;
                const unsigned long long m = __ballot(lane >= first && (double)(x - consumed) >= D);
                if (!m) {
                    s += (float)(total - consumed);
                    break;
                }
                const int j = __ffsll((long long)m) - 1;
                const int before = j ? __builtin_amdgcn_readlane(x, j - 1) : 0;
                const int at = __builtin_amdgcn_readlane(x, j);
                s += (float)(before - consumed);
                s += (float)(at - before);
                consumed = at;
                first = j + 1;
            }
        }
    }
    return s;
}

// GetVarUp (block_match.cpp:233-240) through its short return type
__device__ __forceinline__ float var_up(dme_mv pred, dme_mv m) {
    const int a = norm1(m.x - pred.x, m.y - pred.y), c = norm1(m.x, m.y);
    return (float)(int16_t)(a < c ? a : c);
}

// ---------------------------------------------------------------------------
// Lambda maps on the device: MEData::SetLambdaMap(num_refs, lambda)
// (motion.cpp:128-171) from the pel-search vectors, with FindTransitions'
// vector part (motion.cpp:253-331; its SAD part is overwritten and has no
// effect), then SetLambdaMap(level, map, wt) (motion.cpp:173-199) for the
// level-1 and level-0 maps of ModeDecider.  One workgroup per job.
//
// FindTransitions marks a block when its local vector variation val (an
// integer) reaches thr = mean + 3 sd, computed there in long double and
// rounded to double.  Here the sums are exact (val integer) and thr is
// computed in double, within 1e-12 of the exact value, as is the
// reference's; so `val >= thr` is decided by the integer c = ceil(thr)
// unless thr lies within 1e-9 of an integer.  Such a job is flagged, and the
// host computes its maps by the reference's own arithmetic instead.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_lambda(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                 int *__restrict__ flags) {
    __shared__ unsigned long long s_s1;
    __shared__ double s_q[32];
    __shared__ int s_c, s_amb;
    const Task tk = tasks[blockIdx.x];
    const JobDesc &J = jobs[tk.job];
    const int tid = threadIdx.x, lane = tid & 63, wv = wave_id();
    const int X = J.xnb, Y = J.ynb, N = X * Y;
    if (tid == 0) s_amb = 0;
    for (int r = 0; r < J.nrefs; ++r) {
        const dme_mv *__restrict__ mv = J.vec[r];
        int *__restrict__ tr = J.lwork + (size_t)r * N;
        if (tid == 0) s_s1 = 0;
        __syncthreads();
        unsigned long long s1 = 0;
        for (int k = tid; k < N; k += 1024) {
            const int j = k / X, i = k - j * X;
            int v = 0;
            if (j >= 1 && j < Y - 1 && i >= 1 && i < X - 1) {
                const dme_mv c = mv[k];
#pragma unroll
                for (int q = -1; q <= 1; ++q)
#pragma unroll
                    for (int p = -1; p <= 1; ++p) {
                        const dme_mv n = mv[k + q * X + p];
                        v = max(v, abs(n.x - c.x) + abs(n.y - c.y));
                    }
            }
            tr[k] = v;
            s1 += (unsigned long long)v;
        }
        for (int m = 32; m >= 1; m >>= 1) s1 += __shfl_xor(s1, m, WAVE);
        if (lane == 0) atomicAdd(&s_s1, s1);
        __syncthreads();
        const double mean = (double)s_s1 / (double)N;
        double q = 0.0;
        for (int k = tid; k < N; k += 1024) {
            const int j = k / X, i = k - j * X;
            if (j >= 1 && j < Y - 1 && i >= 1 && i < X - 1) {
                const double d = (double)tr[k] - mean;
                q += d * d;
            }
        }
        for (int m = 32; m >= 1; m >>= 1) q += __shfl_xor(q, m, WAVE);
        if (lane == 0) s_q[wv] = q;
        __syncthreads();
        if (tid == 0) {
            double Q = 0.0;
            for (int w = 0; w < 16; ++w) Q += s_q[w];
            const double thr = mean + 3.0 * sqrt(Q / (double)N);
            const double fl = floor(thr), margin = 1e-9 * (1.0 + thr);
            s_c = (int)ceil(thr);
            if (thr - fl < margin || fl + 1.0 - thr < margin) s_amb = 1;
        }
        __syncthreads();
        const int c = s_c;
        for (int k = tid; k < N; k += 1024) tr[k] = tr[k] >= c ? 1 : 0;
        __syncthreads();
        // a 4 x 4 group of blocks with a transition is marked whole
        const int GX = X / 4, GY = Y / 4;
        for (int g = tid; g < GX * GY; g += 1024) {
            const int gy = g / GX, gx = g - gy * GX;
            int *base = tr + (size_t)(4 * gy) * X + 4 * gx;
            int any = 0;
#pragma unroll
            for (int yy = 0; yy < 4; ++yy)
#pragma unroll
                for (int xx = 0; xx < 4; ++xx) any |= base[yy * X + xx];
#pragma unroll
            for (int yy = 0; yy < 4; ++yy)
#pragma unroll
                for (int xx = 0; xx < 4; ++xx) base[yy * X + xx] = any;
        }
        __syncthreads();
    }
    const float lambda = J.lambda;
    for (int k = tid; k < N; k += 1024) {
        const int j = k / X, i = k - j * X;
        const int t1 = J.lwork[k], t2 = J.nrefs > 1 ? J.lwork[N + k] : 0;
        float m;
        if (J.nrefs == 1) m = t1 ? 0.0f : lambda;
        else if (t1 && t2) m = 0.0f;
        else if (t1 || t2) m = (float)((double)lambda / 4.0);
        else m = lambda;
        if (i < 4 || j < 4) m = (float)((double)m / 5.0);
        J.lam[k] = m;
    }
    __syncthreads();
    for (int level = 1; level >= 0; --level) {
        const int f = 1 << (2 - level), Xl = X / f, Yl = Y / f;
        const float wt = (float)(1.0 / (double)J.lf[level]);
        float *__restrict__ out = level ? J.lam1 : J.lam0;
        for (int k = tid; k < Xl * Yl; k += 1024) {
            const int j = k / Xl, i = k - j * Xl;
            const float *src = J.lam + (size_t)(f * j) * X + f * i;
            float m = src[0];
            for (int qq = 0; qq < f; ++qq)
                for (int pp = 0; pp < f; ++pp) m = fmaxf(src[qq * X + pp], m);
            out[k] = m * wt;
        }
    }
    if (tid == 0) flags[tk.offset] = s_amb;
}

// ---------------------------------------------------------------------------
// Sub-pel refinement, bulk part.  RefineMatchSubp (block_match.cpp:327-467)
// only ever evaluates positions within 3 quarter pels (1 half pel at half
// precision) of the pel vector: stage 1 moves at most one half pel per axis,
// stage 2 one quarter pel more.  Their integer SADs do not depend on the
// prediction, so all (2*3+1)^2 of them are tabulated per block here; the
// wavefront then only adds costs and compares.  One wave per block: the
// picture block and, from the quarter-pel planes, the region of each of the
// 16 phases that the candidates touch are staged in LDS as bytes (one dword
// load brings the four horizontal phases of a position); each lane then
// takes one candidate with packed v_sad_u8.  Record layout (SUB_SLOT ints
// per block, blocks in wavefront order): SAD (dy+rad)*Q + dx+rad, then
// lambda, pel vector and pel costs at 49..54.
// ---------------------------------------------------------------------------
#define SUB_SLOT 56                    // 49 SADs, lambda, pel vector, pel costs
#define SUB_PP 8                       // packed picture row: 32 bytes

// The sub-pel table: candidates within rad units of 1/2^tq pel around the
// scaled integer vector, tq = min(prec, 2) (eighth-pel refinement tabulates
// its half- and quarter-pel stages; its eighth-pel stage is computed on
// demand).
__device__ __forceinline__ int sub_radius(int prec) { return prec >= 2 ? 3 : 1; }
__device__ __forceinline__ int sub_unit(int prec) { return prec < 2 ? prec : 2; }

// Four qp dwords of consecutive positions (each the four horizontal phases
// of one position) to the four phase dwords (each one phase of the four
// positions): a 4 x 4 byte transpose with v_perm.
__device__ __forceinline__ void qp_transpose(const uint32_t d[4], uint32_t e[4]) {
    const uint32_t lo01 = __builtin_amdgcn_perm(d[1], d[0], 0x05010400u), hi01 = __builtin_amdgcn_perm(d[1], d[0], 0x07030602u);
    const uint32_t lo23 = __builtin_amdgcn_perm(d[3], d[2], 0x05010400u), hi23 = __builtin_amdgcn_perm(d[3], d[2], 0x07030602u);
    e[0] = __builtin_amdgcn_perm(lo23, lo01, 0x05040100u);
    e[1] = __builtin_amdgcn_perm(lo23, lo01, 0x07060302u);
    e[2] = __builtin_amdgcn_perm(hi23, hi01, 0x05040100u);
    e[3] = __builtin_amdgcn_perm(hi23, hi01, 0x07060302u);
}

// pp: dwords per picture row (4 * chunks); prows: picture rows; qrows: rows
// of one phase region; qd: dwords per region row
__global__ __launch_bounds__(256) void k_subpel_bulk(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                     int ntasks, int total, int pp, int prows, int qrows, int qd) {
    extern __shared__ uint32_t s_sub[];
    const int wv = wave_id(), lane = threadIdx.x & 63;
    const int gi = xcd_block() * 4 + wv;
    if (gi >= total) return;  // each wave works alone
    const int plane = qrows * qd;  // dwords of one phase region
    uint32_t *s_pic = s_sub + wv * (pp * prows + 16 * plane);
    uint32_t *s_q = s_pic + pp * prows;  // [fy * 4 + fx][row][qd]
    const Task tk = tasks[find_task_wave(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    const int r = tk.ref, k = gi - tk.offset;
    const int X = J.xnb, W = J.W, H = J.H, QW = W + 2;
    const int xb = k % X, yb = k / X;
    const int prec = J.prec, rad = sub_radius(prec), Q = 2 * rad + 1, QS = Q * Q;
    const Blk b = block_limits(J.blen[2], J.bsep[2], J.boff[2], W, H, xb, yb);
    const dme_mv pel = J.vec[r][k];
    int *__restrict__ rec = J.qtab[r] + (size_t)wf_pos(xb, yb, X, J.ynb) * SUB_SLOT;
    if (!b.empty() && prec > 0) {
        // candidate range in quarter pels and the integer offsets it spans
        const int tq = sub_unit(prec), sq = 2 - tq;  // quarter pels per candidate unit
        const int cqx = (pel.x << tq) << sq, cqy = (pel.y << tq) << sq, rq = rad << sq;
        const int qx0 = (cqx - rq) >> 2, qy0 = (cqy - rq) >> 2;
        const int RW = b.xl + ((cqx + rq) >> 2) - qx0, RH = b.yl + ((cqy + rq) >> 2) - qy0;
        // picture block from the byte plane; its loads go out with the first
        // phase-region loads
        const int ppad = (b.xl + 15) & ~15;
        const B8Rect pr = b8_rect(J.pic8, J.p8, s_pic, pp, b.xp, b.yp, b.xl, b.yl, ppad);
        const bool pfast = b8_inside(J.p8, H, b.xp, b.yp, b.yl, ppad);
        uint32_t plo[4], phi[4];
        if (pfast) b8_load(pr, J.p8, 0, plo, phi);  // at most 32 rows, 8 or more per pass
        // phase regions from the quarter-pel planes: element (R, c) holds
        // positions X0 + 4c .. + 3 of qp row R = ry * 4 + fy (position row
        // Y0 + ry clamped to [-1, H], phase row fy), four dword loads
        // transposed into one dword of each of the four phase planes fx
        const uint32_t *__restrict__ qp = J.qp[r];
        const int X0 = b.xp + qx0, Y0 = b.yp + qy0;
        const int ncd = (RW + 3) >> 2, rpp = WAVE / ncd;
        const int lr = (int)(((float)lane + 0.5f) * (1.0f / (float)ncd)), lc = lane - lr * ncd;
        const bool on = lr < rpp;
        int xo[4];  // qp column (position + 1) of this lane's four positions, clamped to [-1, W]
#pragma unroll
        for (int i = 0; i < 4; ++i) xo[i] = clampi(X0 + 4 * lc + i, -1, W) + 1;
        const int NR = 4 * RH, np = (NR + rpp - 1) / rpp;
        for (int p0 = 0; p0 < np; p0 += 4) {
            uint32_t d[4][4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int R = lr + (p0 + e) * rpp;
#pragma unroll
                for (int i = 0; i < 4; ++i) d[e][i] = 0;
                if (on && R < NR) {
                    const uint32_t *row = qp + (size_t)((clampi(Y0 + (R >> 2), -1, H) + 1) * 4 + (R & 3)) * QW;
#pragma unroll
                    for (int i = 0; i < 4; ++i) d[e][i] = row[xo[i]];
                }
            }
            if (p0 == 0) {
                if (pfast) b8_store(pr, 0, plo, phi);
                else b8_copy_clamped(pr, J.pic8, J.p8, W, H, b.xp, b.yp);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int R = lr + (p0 + e) * rpp;
                if (on && R < NR) {
                    uint32_t t[4];
                    qp_transpose(d[e], t);
                    uint32_t *dst = s_q + (R & 3) * 4 * plane + (R >> 2) * qd + lc;
#pragma unroll
                    for (int fx = 0; fx < 4; ++fx) dst[fx * plane] = t[fx];
                }
            }
        }
        WAVE_SYNC();
        if (lane < QS) {
            const int mx = cqx + ((lane % Q - rad) << sq), my = cqy + ((lane / Q - rad) << sq);
            const int fx = mx & 3, fy = my & 3;
            const uint32_t *src = s_q + (fy * 4 + fx) * plane;
            const int bo = (mx >> 2) - qx0, r0 = (my >> 2) - qy0;
            rec[lane] = (b.xl == 16 && b.yl == 16 && pp == 4 && qd == 5)
                            ? sad16x16_u8<4, 5>(s_pic, src, bo, r0)
                            : ((b.xl & 15) == 0 ? packed_sad_u8<false>(s_pic, pp, src, qd, bo, r0, b.xl, b.yl)
                                                : packed_sad_u8<true>(s_pic, pp, src, qd, bo, r0, b.xl, b.yl));
        }
    }
    if (lane == 0) {
        rec[49] = __float_as_int(J.lam[k]);
        rec[50] = pel.x;
        rec[51] = pel.y;
        const dme_cost pc = J.pc[r][k];
        rec[52] = __float_as_int(pc.SAD);
        rec[53] = __float_as_int(pc.mvcost);
        rec[54] = __float_as_int(pc.total);
    }
}

// float total of a bail-out Diff (me_utils.cpp:481-571, 686-867): MMX
// semantics add the integer SAD once to the start cost in bounds
// (me_utils_mmx.cpp:270-278); the bounds-checked path and the plain C build
// add per pixel in float.  That per-pixel sum equals the single addition
// whenever every partial sum is representable, which holds when the start is
// a multiple of the float spacing at the final magnitude; otherwise it is
// replayed pixel by pixel.
// With m = start + isad in [2^(e-1), 2^e), the float spacing there is
// u = 2^(e-24): the test is u <= 1 and start a multiple of u, read off the
// exponent of m and the lowest set bit of start's significand.
__device__ __forceinline__ bool chain_is_exact(float start, int isad) {
    const double m = (double)start + (double)isad;
    if (m < 1.0) return start == 0.0f;
    const int e = (int)((__double_as_longlong(m) >> 52) & 0x7ff) - 1022;
    if (e > 24) return false;
    if (start == 0.0f) return true;
    const unsigned bits = __float_as_uint(start);
    const int ex = (int)((bits >> 23) & 0xffu);
    const unsigned mant = (bits & 0x7fffffu) | (ex ? 0x800000u : 0u);
    return (ex ? ex : 1) - 150 + __builtin_ctz(mant) >= e - 24;  // start = mant * 2^((ex ? ex : 1) - 150)
}


// ---------------------------------------------------------------------------
// Sub-pel refinement, wavefront part (SubpelRefine::MatchPic/DoBlock/GetPred,
// me_subpel.cpp:101-176; BlockMatcher::RefineMatchSubp,
// block_match.cpp:327-467).  One thread per block row, wavefront t = x + 2y;
// the left, top-left, top and top-right refined vectors come from a 4-slot
// LDS ring per row, and the workgroup copies the records of diagonal t + 1
// into the other LDS bank while step t runs.  Candidate SADs come from the
// record's table.  What the table cannot answer is queued in LDS and taken
// by the workgroup's waves, one entry per wave: predictor SADs outside the
// table, and totals that need the per-pixel float chain.  The four neighbour
// tests of a stage do not depend on each other's outcome, so their chains are
// queued together, then those of the up-to-two second-round tests.
// ---------------------------------------------------------------------------
struct SubJob {
    int xp, yp, xl, yl;  // block
    int mx, my, step;    // vector and its units (1: half pel, 2: quarter pel)
    int ires;            // integer SAD (predictor SADs)
    float start, fres;   // chain start and result
    int pad[2];
};
#define SUB_QCAP 256

// One queue round trip: every thread with want set gets a slot, the waves
// serve the entries, each thread gets its result back.  Uniform call.  Round
// trips alternate between two counters: a counter is read after the first
// barrier of its round trip and reset (when it was used) before the next
// round trip's first barrier, which every thread passes before the round trip
// after that touches the counter again; an empty round trip costs one barrier.
template <bool CHAIN, bool E8, bool COMB>
__device__ __forceinline__ void sub_serve(int *cnts, int &phase, SubJob *queue, const JobDesc &J,
                                          const uint32_t *__restrict__ qp, const SubJob (&mine)[4],
                                          const bool (&want)[4], int (&ires)[4], float (&fres)[4]) {
    const int tid = threadIdx.x, nth = blockDim.x, lane = tid & 63, wv = tid >> 6, nwv = nth >> 6;
    int *cnt = cnts + phase;
    phase ^= 1;
    int slot[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) slot[j] = want[j] ? atomicAdd(cnt, 1) : -1;
    __syncthreads();
    const int nq = *cnt;
    if (nq == 0) return;  // uniform
    for (int m0 = 0; m0 < nq; m0 += SUB_QCAP) {
        const int nm = min(SUB_QCAP, nq - m0);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (slot[j] >= m0 && slot[j] < m0 + nm) queue[slot[j] - m0] = mine[j];
        __syncthreads();
        const int UW = 2 * J.W, UH = 2 * J.H;
        for (int e = wv; e < nm; e += nwv) {
            const SubJob q = queue[e];
            Blk b;
            b.xp = q.xp;
            b.yp = q.yp;
            b.xl = q.xl;
            b.yl = q.yl;
            dme_mv m;
            m.x = q.mx;
            m.y = q.my;
            if (CHAIN) {
                const UpStart us = up_start<E8>(b, m, q.step, UW, UH);
                const float v = wave_chain<E8, COMB>(q.start, J.pic[0], J.W, qp, UW, UH, b, us);
                if (lane == 0) queue[e].fres = v;
            } else {
                int v = 0;
                wave_up_sads<E8, COMB>(J.pic[0], J.W, qp, UW, UH, b, q.step, &m, 1, &v);
                if (lane == 0) queue[e].ires = v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (slot[j] >= m0 && slot[j] < m0 + nm) {
                ires[j] = queue[slot[j] - m0].ires;
                fres[j] = queue[slot[j] - m0].fres;
            }
        __syncthreads();
    }
    if (tid == 0) *cnt = 0;
}

// E8: the eighth-pel jobs of the launch (the others exit), else the rest;
// COMB likewise for the combined-component jobs.
template <bool E8, bool COMB>
__global__ __launch_bounds__(512) void k_subpel_wave(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                      int max_rows) {
    extern __shared__ int s_dyn[];  // ring [max_rows][4] mv | slots [2][max_rows][SUB_SLOT] | queue [SUB_QCAP]
    __shared__ int s_cnt[2];
    __shared__ int s_chain[2];  // per pass parity: some block of the pass may need a float chain
    const Task tk = tasks[blockIdx.x];
    const JobDesc &J = jobs[tk.job];
    if (tk.ref >= J.nrefs || J.prec == 0 || (J.prec == 3) != E8 || (J.comb != 0) != COMB) return;  // uniform
    const int r = tk.ref, prec = J.prec, rad = sub_radius(prec), Q = 2 * rad + 1;
    const int X = J.xnb, Y = J.ynb, W = J.W, H = J.H;
    const int UW = 2 * W, UH = 2 * H;
    dme_mv *ring = reinterpret_cast<dme_mv *>(s_dyn);
    int *slots = s_dyn + 8 * max_rows;
    SubJob *queue = reinterpret_cast<SubJob *>(slots + 2 * max_rows * SUB_SLOT);
    const uint32_t *__restrict__ qp = J.qp[r];
    const int *__restrict__ qtab = J.qtab[r];
    dme_mv *__restrict__ V = J.vec[r];
    dme_cost *__restrict__ C = J.pc[r];
    int blen[2] = {J.blen[2][0], J.blen[2][1]}, bsep[2] = {J.bsep[2][0], J.bsep[2][1]};
    int boff[2] = {J.boff[2][0], J.boff[2][1]};
    const bool c_sem = J.fp_sem == DME_FP_C;
    if (threadIdx.x < 2) s_cnt[threadIdx.x] = s_chain[threadIdx.x] = 0;
    int phase = 0, gen = 0;
    const int T = (X - 1) + 2 * (Y - 1);
    for (int t = -1; t <= T; ++t) {
        WfCopy<8> cp;
        wf_issue(cp, qtab, slots, SUB_SLOT, max_rows, t + 1, X, Y);
        for (int base = 0; base < Y; base += blockDim.x) {
            __syncthreads();  // ring, bank and counter writes are visible
            const int yb = base + threadIdx.x;
            const int xb = t - 2 * yb;
            const bool active = yb < Y && xb >= 0 && xb < X;
            const int *mytab = slots + ((xb & 1) * max_rows + (active ? yb : 0)) * SUB_SLOT;
            dme_mv pred, pel;
            pred.x = pred.y = pel.x = pel.y = 0;
            dme_cost cost;
            cost.SAD = cost.mvcost = cost.total = 0.0f;
            float lambda = 0.0f, lim = 0.0f;
            Blk b;
            b.xp = b.yp = b.xl = b.yl = 0;
            bool refine = false, miss = false;
            dme_mv out_mv;
            dme_cost out_cost;
            out_mv.x = out_mv.y = 0;
            out_cost = cost;
            int psad = 0;
            if (active) {
                // GetPred: median of the available left, top-left, top, top-right
                dme_mv nb0, nb1, nb2, nb3;
                nb0 = nb1 = nb2 = nb3 = pred;
                int nn = 0;
                auto push = [&](dme_mv v) {
                    if (nn == 0) nb0 = v;
                    else if (nn == 1) nb1 = v;
                    else if (nn == 2) nb2 = v;
                    else nb3 = v;
                    ++nn;
                };
                if (xb > 0) push(ring[yb * 4 + (t - 1) % 4]);
                if (xb > 0 && yb > 0) push(ring[(yb - 1) * 4 + (t - 3) % 4]);
                if (yb > 0) push(ring[(yb - 1) * 4 + (t - 2) % 4]);
                if (yb > 0 && xb + 1 < X) push(ring[(yb - 1) * 4 + (t - 1) % 4]);
                if (nn == 1) pred = nb0;
                else if (nn == 2) pred = mean2(nb0, nb1);
                else if (nn == 3) pred = med3(nb0, nb1, nb2);
                else if (nn == 4) {
                    const int sx = nb0.x + nb1.x + nb2.x + nb3.x, sy = nb0.y + nb1.y + nb2.y + nb3.y;
                    const int mxx = max(max(nb0.x, nb1.x), max(nb2.x, nb3.x));
                    const int mnx = min(min(nb0.x, nb1.x), min(nb2.x, nb3.x));
                    const int mxy = max(max(nb0.y, nb1.y), max(nb2.y, nb3.y));
                    const int mny = min(min(nb0.y, nb1.y), min(nb2.y, nb3.y));
                    pred.x = (sx - (mxx + mnx) + 1) >> 1;
                    pred.y = (sy - (mxy + mny) + 1) >> 1;
                }
                lambda = __int_as_float(mytab[49]);
                b = block_limits(blen, bsep, boff, W, H, xb, yb);
                pel.x = mytab[50];
                pel.y = mytab[51];
                cost.SAD = __int_as_float(mytab[52]);
                cost.mvcost = __int_as_float(mytab[53]);
                cost.total = __int_as_float(mytab[54]);
                dme_mv pel_up;
                pel_up.x = pel.x << prec;
                pel_up.y = pel.y << prec;
                cost.mvcost = var_up(pred, pel_up);
                cost.total = cost.SAD + lambda * cost.mvcost;
                lim = (float)(2 * b.xl * b.yl);
                if (cost.SAD < lim) {
                    out_mv = pel_up;
                    out_cost = cost;
                } else {
                    refine = true;
                    if (!b.empty()) {
                        // an eighth-pel predictor is always computed on demand
                        const int dx = pred.x - pel_up.x, dy = pred.y - pel_up.y;
                        if ((!E8 || prec <= 2) && abs(dx) <= rad && abs(dy) <= rad)
                            psad = mytab[(dy + rad) * Q + dx + rad];
                        else miss = true;
                    }
                }
            }
            // Can any candidate this block may test need the per-pixel float
            // chain?  Only where some candidate's samples leave the picture
            // (the window's extreme candidates bound that) or with C
            // semantics, and only when a start cost mvcost * lambda can fall
            // off the float grid at the total's magnitude: mvcost is an
            // integer below N = |window|_1 (< 32768), so the start is a
            // multiple of lambda's lowest set bit, and totals stay below
            // N * lambda + 255 * xl * yl.  When no block of the pass can, its
            // chain round trips are skipped.
            const int cf = gen & 1;
            if (threadIdx.x == 0) s_chain[cf ^ 1] = 0;  // the previous pass read it before this pass's barrier
            if (refine && !b.empty()) {
                const int cqx = pel.x << prec, cqy = pel.y << prec;
                bool edge = c_sem;
                for (int i = 1; i <= prec && !edge; ++i) {
                    const int sh = prec - i, r1 = rad >> sh;
                    dme_mv lo, hi;
                    lo.x = (cqx >> sh) - r1;
                    lo.y = (cqy >> sh) - r1;
                    hi.x = (cqx >> sh) + r1;
                    hi.y = (cqy >> sh) + r1;
                    edge = up_start<E8>(b, lo, i, UW, UH).bounds || up_start<E8>(b, hi, i, UW, UH).bounds;
                }
                bool may = false;
                if (edge && lambda != 0.0f) {
                    const int N = abs(cqx) + abs(cqy) + 2 * rad;
                    const unsigned lb = __float_as_uint(lambda);
                    const int lex = (int)((lb >> 23) & 0xffu);
                    const unsigned lm = (lb & 0x7fffffu) | (lex ? 0x800000u : 0u);
                    const int low = (lex ? lex : 1) - 150 + __builtin_ctz(lm);  // lowest set bit of lambda
                    const double mmax = (double)N * fabs((double)lambda) + 255.0 * b.xl * b.yl;
                    const int e = (int)((__double_as_longlong(mmax) >> 52) & 0x7ff) - 1022;  // mmax < 2^e
                    may = N >= 32768 || lambda < 0.0f || low < e - 24;
                }
                if (may || (E8 && prec == 3)) s_chain[cf] = 1;  // the eighth-pel stage always adds per pixel
            }
            // predictor SADs outside the table
            {
                SubJob mine[4];
                bool want[4] = {miss, false, false, false};
                int ir[4] = {psad, 0, 0, 0};
                float fr[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                mine[0].xp = b.xp;
                mine[0].yp = b.yp;
                mine[0].xl = b.xl;
                mine[0].yl = b.yl;
                mine[0].mx = pred.x;
                mine[0].my = pred.y;
                mine[0].step = prec;
                mine[0].ires = 0;
                mine[0].start = mine[0].fres = 0.0f;
                mine[1] = mine[2] = mine[3] = mine[0];
                sub_serve<false, E8, COMB>(s_cnt, phase, queue, J, qp, mine, want, ir, fr);
                psad = ir[0];
            }
            const bool chains = s_chain[cf] != 0;  // visible after the round trip's barrier
            ++gen;
            dme_cost pcost;
            pcost.mvcost = 0.0f;
            pcost.SAD = (float)psad;
            pcost.total = pcost.SAD;
            if (refine && pcost.SAD < lim) {
                out_mv = pred;
                out_cost = pcost;
                refine = false;
            }
            // RefineMatchSubp stages; every thread takes part in the queue
            // round trips, the refining ones with their candidates
            dme_cost best = cost;
            dme_mv bmv = pel;
            bool live = refine, bailed = false;
            const bool tests = !b.empty();
            const int tq = E8 ? sub_unit(prec) : prec;  // the table's unit: 1/2^tq pel
            const int cx = pel.x << tq, cy = pel.y << tq;
            for (int i = 1; i <= prec; ++i) {
                if (live) {
                    bmv.x <<= 1;
                    bmv.y <<= 1;
                }
                dme_mv tb = bmv;
                const int sh = prec - i;
                for (int round = 0; round < 2; ++round) {
                    dme_mv m[4];
                    bool on[4];
                    if (round == 0) {
                        m[0].x = bmv.x - 1; m[0].y = bmv.y;
                        m[1].x = bmv.x + 1; m[1].y = bmv.y;
                        m[2].x = bmv.x; m[2].y = bmv.y - 1;
                        m[3].x = bmv.x; m[3].y = bmv.y + 1;
#pragma unroll
                        for (int j = 0; j < 4; ++j) on[j] = live && tests;
                    } else {
                        const bool mx = tb.x != bmv.x, my = !mx && tb.y != bmv.y;
                        m[0].x = mx ? tb.x : tb.x - 1; m[0].y = mx ? tb.y - 1 : tb.y;
                        m[1].x = mx ? tb.x : tb.x + 1; m[1].y = mx ? tb.y + 1 : tb.y;
                        m[2] = m[3] = tb;
                        on[0] = on[1] = live && tests && (mx || my);
                        on[2] = on[3] = false;
                    }
                    int sd[4];
                    float mvc[4], st[4], tot[4];
                    bool need[4];
                    SubJob mine[4];
                    // stages in half and quarter pel read the table; the
                    // eighth-pel stage (BlockDiffEighthPel, per-pixel float
                    // adds in both builds) is a chain round trip
                    const bool tab = !E8 || i <= tq;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        dme_mv mf;
                        mf.x = m[j].x << sh;
                        mf.y = m[j].y << sh;
                        const int ti = tab ? ((m[j].y << (tq - i)) - cy + rad) * Q + ((m[j].x << (tq - i)) - cx + rad) : 0;
                        sd[j] = on[j] && tab ? mytab[ti] : 0;
                        mvc[j] = var_up(pred, mf);
                        st[j] = mvc[j] * lambda;
                        tot[j] = (float)sd[j] + st[j];
                        const UpStart us = up_start<E8>(b, m[j], i, UW, UH);
                        need[j] = on[j] && (!tab || ((us.bounds || c_sem) && !chain_is_exact(st[j], sd[j])));
                        mine[j].xp = b.xp;
                        mine[j].yp = b.yp;
                        mine[j].xl = b.xl;
                        mine[j].yl = b.yl;
                        mine[j].mx = m[j].x;
                        mine[j].my = m[j].y;
                        mine[j].step = i;
                        mine[j].ires = 0;
                        mine[j].start = st[j];
                        mine[j].fres = 0.0f;
                    }
                    int ir[4] = {0, 0, 0, 0};
                    float fr[4] = {tot[0], tot[1], tot[2], tot[3]};
                    if (chains) sub_serve<true, E8, COMB>(s_cnt, phase, queue, J, qp, mine, need, ir, fr);  // uniform
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (on[j] && fr[j] < best.total) {
                            tb = m[j];
                            best.total = fr[j];
                            best.mvcost = mvc[j];
                            best.SAD = fr[j] - st[j];
                        }
                    }
                }
                if (live) {
                    bmv = tb;
                    if ((double)best.total > 1.1 * (double)pcost.total) {
                        bailed = true;
                        live = false;
                    }
                }
            }
            if (refine) {
                if (bailed) {
                    out_mv = pred;
                    out_cost = pcost;
                } else {
                    out_mv = bmv;
                    out_cost = best;
                }
            }
            if (active) {
                const size_t k = (size_t)yb * X + xb;
                V[k] = out_mv;
                C[k] = out_cost;
                ring[yb * 4 + t % 4] = out_mv;
            }
        }
        wf_finish(cp);
        __syncthreads();
    }
}

// Pixel-precision path: vectors doubled before mode decision and halved
// after it (picture_compress.cpp:240-259, 274-296).
__global__ __launch_bounds__(256) void k_shift_vectors(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                       int ntasks, int total, int left) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= total) return;
    const Task tk = tasks[find_task(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    dme_mv &v = J.vec[tk.ref][gi - tk.offset];
    if (left) { v.x <<= 1; v.y <<= 1; }
    else { v.x >>= 1; v.y >>= 1; }
}

// ---------------------------------------------------------------------------
// Mode decision helpers
// ---------------------------------------------------------------------------

// The bi-prediction SAD of a block from the fast form's byte up-planes
// (defined with them below).
__device__ int up8_bi_sad(const JobDesc &J, const Blk &b, dme_mv m1, dme_mv m2);

// The second reference's start of BiBlockEighthPel: its own, unless its
// window takes the bounds-checked path, which reads it from the first
// reference's start with its own remainders (me_utils.cpp:1764-1768).
__device__ __forceinline__ UpStart bi_second(const UpStart &s1, UpStart s2) {
    if (s2.e && s2.bounds) {
        s2.rsx = s1.rsx;
        s2.rsy = s1.rsy;
    }
    return s2;
}

// BiBlock{Half,Quarter,Eighth}Pel::Diff (me_utils.cpp:1172-1785): sum of
// |((2 pic - r1) - r2) >> 1| with short intermediates (the eighth-pel
// form's float adds of integers are exact).
template <bool E8, bool COMB, int NB = 8>
__device__ int wave_bi_sad(const JobDesc &J, const Blk &b, dme_mv m1, dme_mv m2) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int UW = 2 * J.W, UH = 2 * J.H;
    const UpStart s1 = up_start<E8>(b, m1, J.step, UW, UH);
    const UpStart s2 = E8 ? bi_second(s1, up_start<E8>(b, m2, J.step, UW, UH)) : up_start<E8>(b, m2, J.step, UW, UH);
    const int16_t *__restrict__ pic = J.pic[0];
    const uint32_t *__restrict__ q1 = J.qp[0];
    const uint32_t *__restrict__ q2 = J.qp[1];
    int s = 0;
    const int n = b.xl * b.yl;
    for (int b0 = 0; b0 < n; b0 += NB * WAVE) {
        int p[NB], r1[NB], r2[NB];
        bool ok[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int i = b0 + k * WAVE + lane;
            ok[k] = i < n;
            const int ii = ok[k] ? i : 0;
            const int dy = ii / b.xl, dx = ii - dy * b.xl;
            p[k] = pic[(unsigned)((b.yp + dy) * J.W + b.xp + dx)];
            r1[k] = qp_at<E8, COMB>(q1, J.W, J.H, s1, dy, dx);
            r2[k] = qp_at<E8, COMB>(q2, J.W, J.H, s2, dy, dx);
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int16_t diff = (int16_t)((p[k] << 1) - r1[k]);
            const int16_t temp = (int16_t)(((int)diff - r2[k]) >> 1);
            s += ok[k] ? abs((int)temp) : 0;
        }
    }
    return wave_sum(s);
}

// The same sum for half- and quarter-pel luma jobs from the byte planes,
// four samples per lane-op: a quarter-pel window's samples share one phase
// (fx, fy), so each dword of four is picked out of four quarter-pel-plane
// dwords (one per position), and |((2 p - r1) - r2) >> 1| =
// |p - ceil((r1 + r2) / 2)| is a v_sad_u8 against the bytewise rounded-up
// average (the bias of 128 cancels).  Lane (row, dword column) items.
#ifndef BI_U8_ITEMS
#define BI_U8_ITEMS 2
#endif
__device__ __forceinline__ uint32_t qp_phase_dword(const uint32_t *__restrict__ qp, int W, int H, int X, int Y, int fy,
                                                   uint32_t sel) {
    const uint32_t *__restrict__ qr = qp + (size_t)((clampi(Y, -1, H) + 1) * 4 + fy) * (W + 2);
    uint32_t d[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) d[i] = qr[clampi(X + i, -1, W) + 1];
    const uint32_t lo = __builtin_amdgcn_perm(d[1], d[0], sel), hi = __builtin_amdgcn_perm(d[3], d[2], sel);
    return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

__device__ __forceinline__ uint32_t avg_up_u8(uint32_t a, uint32_t b);
__device__ __forceinline__ uint32_t b8_dword(const uint8_t *__restrict__ origin, int pitch, int x, int y);

template <int U = BI_U8_ITEMS>
__device__ int wave_bi_sad_u8(const JobDesc &J, const Blk &b, dme_mv m1, dme_mv m2) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int W = J.W, H = J.H;
    const UpStart s1 = up_start<false>(b, m1, J.step, 2 * W, 2 * H), s2 = up_start<false>(b, m2, J.step, 2 * W, 2 * H);
    const int xq1 = 2 * s1.rsx + s1.rx, yq1 = 2 * s1.rsy + s1.ry, xq2 = 2 * s2.rsx + s2.rx, yq2 = 2 * s2.rsy + s2.ry;
    const uint32_t sel1 = (uint32_t)(xq1 & 3) | ((uint32_t)((xq1 & 3) + 4) << 8) | 0x0c0c0000u;
    const uint32_t sel2 = (uint32_t)(xq2 & 3) | ((uint32_t)((xq2 & 3) + 4) << 8) | 0x0c0c0000u;
    const int nd = (b.xl + 3) >> 2, items = nd * b.yl;
    uint32_t acc = 0;
    // U items per lane with all their loads in flight together
    for (int i0 = 0; i0 < items; i0 += U * WAVE) {
        uint32_t p[U], r1[U], r2[U], mask[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int it = i0 + u * WAVE + lane;
            p[u] = r1[u] = r2[u] = mask[u] = 0;
            if (it < items) {
                const int row = it / nd, col = it - row * nd;
                const int valid = b.xl - 4 * col;
                mask[u] = valid >= 4 ? 0xffffffffu : ((1u << (8 * valid)) - 1u);
                p[u] = b8_dword(J.pic8, J.p8, b.xp + 4 * col, b.yp + row);
                r1[u] = qp_phase_dword(J.qp[0], W, H, (xq1 >> 2) + 4 * col, (yq1 >> 2) + row, yq1 & 3, sel1);
                r2[u] = qp_phase_dword(J.qp[1], W, H, (xq2 >> 2) + 4 * col, (yq2 >> 2) + row, yq2 & 3, sel2);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc = __builtin_amdgcn_sad_u8(p[u] & mask[u], avg_up_u8(r1[u], r2[u]) & mask[u], acc);
    }
    return wave_sum((int)acc);
}

// IntraBlockDiff::CalcDC (me_utils.cpp:284-299): truncating integer mean.
__device__ int wave_block_dc(const int16_t *__restrict__ p, int W, const Blk &b) {
    const int lane = threadIdx.x & (WAVE - 1);
    int s = 0;
    const int n = b.xl * b.yl;
    for (int i = lane; i < n; i += WAVE) {
        const int dy = i / b.xl, dx = i - dy * b.xl;
        s += p[(size_t)(b.yp + dy) * W + b.xp + dx];
    }
    s = wave_sum(s);
    return (int16_t)(s / n);
}

__device__ int wave_intra_sad(const int16_t *__restrict__ p, int W, const Blk &b, int dc) {
    const int lane = threadIdx.x & (WAVE - 1);
    int s = 0;
    const int n = b.xl * b.yl;
    for (int i = lane; i < n; i += WAVE) {
        const int dy = i / b.xl, dx = i - dy * b.xl;
        s += abs((int)p[(size_t)(b.yp + dy) * W + b.xp + dx] - dc);
    }
    return wave_sum(s);
}

// DoUnitDecn at level 2 (me_mode_decn.cpp:368-462), for every block in
// parallel: it only reads the block's own refined vectors and costs.  UP8:
// the references are byte up-planes (the learned path's fast form).  E8:
// the eighth-pel jobs of the launch (the others are skipped), else the rest;
// COMB likewise for the combined-component jobs.
template <bool UP8, bool E8, bool COMB>
__global__ __launch_bounds__(256) void k_mode_l2(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                 int ntasks, int total) {
    const int g0 = (xcd_block() * 4 + wave_id()) * PW_MODE_L2, g1 = min(g0 + PW_MODE_L2, total);
    for (int gi = g0; gi < g1; ++gi) {
        WAVE_SYNC();  // the previous block no longer reads the wave's LDS
        const Task tk = tasks[find_task_wave(tasks, ntasks, gi)];
        const JobDesc &J = jobs[tk.job];
        if ((J.step == 3) != E8 || (J.comb != 0) != COMB) continue;  // uniform per wave
        const int k = gi - tk.offset;
        const int xb = k % J.xnb, yb = k / J.xnb;
        const int lane = threadIdx.x & 63;
        const Blk b = block_limits(J.blen[2], J.bsep[2], J.boff[2], J.W, J.H, xb, yb);
        const float loc_lambda = J.lam[k];
        const float lf = J.lf[2];
        const float mode_cost = 0.0f;
        int mode = DME_REF1_ONLY;
        dme_cost c1 = J.pc[0][k];
        c1.total *= lf;
        float min_cost = c1.total + mode_cost;
        float best_sad = c1.SAD;
        if (J.nrefs > 1) {
            dme_cost c2 = J.pc[1][k];
            c2.total *= lf;
            float uc = c2.total + mode_cost;
            if (uc < min_cost) {
                mode = DME_REF2_ONLY;
                min_cost = uc;
                best_sad = c2.SAD;
            }
            dme_cost bc;
            bc.mvcost = c1.mvcost + c2.mvcost;
            bc.SAD = b.empty() ? 0.0f
                               : (float)(UP8 ? up8_bi_sad(J, b, J.vec[0][k], J.vec[1][k])
                                             : (E8 || COMB ? wave_bi_sad<E8, COMB>(J, b, J.vec[0][k], J.vec[1][k])
                                                           : wave_bi_sad_u8(J, b, J.vec[0][k], J.vec[1][k])));
            bc.total = bc.SAD + loc_lambda * bc.mvcost;
            bc.total *= lf;
            uc = bc.total + mode_cost;
            if (uc < min_cost) {
                mode = DME_REF1AND2;
                min_cost = uc;
                best_sad = bc.SAD;
            }
            if (lane == 0) J.bi[k] = bc;
        } else if (lane == 0) {
            // the second reference's arrays stay zero (motion.cpp:55-126)
            dme_cost z;
            z.SAD = z.mvcost = z.total = 0.0f;
            dme_mv zm;
            zm.x = zm.y = 0;
            J.bi[k] = z;
            J.pc[1][k] = z;
            J.vec[1][k] = zm;
        }
        float intra = 0.0f;
        int dc0 = 0;
        if ((double)best_sad > 4.0 * J.blen[2][0] * J.blen[2][1]) {
            int dc = 0;
            float ic = 0.0f;
            if (!b.empty()) {
                dc = wave_block_dc(J.pic[0], J.W, b);
                ic = (float)wave_intra_sad(J.pic[0], J.W, b, dc);
            }
            ic *= lf;
            const float uc = ic + mode_cost;
            if (uc < min_cost && (double)ic < 0.85 * (double)best_sad) {
                mode = DME_INTRA;
                min_cost = uc;
            }
            intra = ic;
            dc0 = dc;
        }
        if (lane == 0) {
            // intra cost and DCs are zero unless set here (motion.cpp:55-126)
            J.intra[k] = intra;
            J.dc[0][k] = (int16_t)dc0;
            J.dc[1][k] = 0;
            J.dc[2][k] = 0;
            J.mode[k] = mode;
            J.ucost[k] = min_cost;
        }
    }
}

// Prediction of ModeDecider::DoME from the level-2 vectors
// (me_mode_decn.cpp:308-321).
__device__ __forceinline__ dme_mv mode_pred(const dme_mv *__restrict__ V2, int X2, int xb, int yb) {
    dme_mv p;
    const size_t k = (size_t)yb * X2 + xb;
    if (xb > 0 && yb > 0) p = med3(V2[k - 1], V2[k - X2 - 1], V2[k - X2]);
    else if (xb == 0 && yb > 0) p = mean2(V2[k - X2], V2[k - X2 + 1]);
    else if (xb > 0 && yb == 0) p = mean2(V2[k - 1], V2[k + X2 - 1]);
    else { p.x = 0; p.y = 0; }
    return p;
}

// ---------------------------------------------------------------------------
// Mode decision tables.  The candidates of a sub-superblock (level 1) are the
// level-2 vectors of its four blocks, fixed before any superblock decision
// (me_mode_decn.cpp:296-300); those of a superblock (level 0) are level-1
// results, hence level-2 vectors of its sixteen blocks.  Their SADs, and the
// bi-prediction SADs of every level-1 candidate pair, do not depend on the
// predictions, so they are computed here in parallel:
//   m1tab[unit * 24 + r*4 + q]      unit candidate q of reference r
//   m1tab[unit * 24 + 8 + q1*4 + q2] bi-prediction of candidates q1, q2
//   m0tab[sb * 32 + r*16 + q]       superblock candidate q (block q of the SB)
// Duplicate vectors are evaluated once.  One wave per unit: the unit's
// picture samples are staged in LDS from the byte plane; for a candidate all
// samples share one quarter-pel phase (fx, fy), so a lane gathers byte fx of
// four consecutive qp dwords of phase row fy with v_perm into one packed
// dword and adds it with v_sad_u8.  The bi-prediction term
// |(2p - r0 - r1) >> 1| (me_utils.cpp BiBlockDiff) equals |p - ceil((r0 + r1)
// / 2)|, a byte SAD against the rounded-up average of the two packed
// candidates, which level 1 keeps in LDS for the pairs.
// ---------------------------------------------------------------------------

// Packed rounded-up average of four byte pairs: (a | b) - ((a ^ b) >> 1).
__device__ __forceinline__ uint32_t avg_up_u8(uint32_t a, uint32_t b) { return (a | b) - (((a ^ b) & 0xfefefefeu) >> 1); }

// Per-launch LDS layout of k_mode_tables, in dwords per wave: picture rows of
// pd dwords (prows of them), then for level 1 eight candidate slots of
// `slot` dwords, rows of spd dwords.
struct MtLds {
    int pd, prows, spd, slot, wave;
};

// E8: the eighth-pel jobs of the launch (their gathers need registers the
// others do not; the other jobs are skipped), else the rest.
template <bool E8>
__global__ __launch_bounds__(256) void k_mode_tables(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                     int ntasks, int total, MtLds ly) {
    extern __shared__ uint32_t s_mt[];
    __shared__ int s_pair[4][16];
    const int wv = wave_id(), lane = threadIdx.x & 63;
    const int g0 = (xcd_block() * 4 + wv) * PW_MODE_TABLES, g1 = min(g0 + PW_MODE_TABLES, total);
    for (int gi = g0; gi < g1; ++gi) {
        WAVE_SYNC();  // the previous block no longer reads the wave's LDS
        const Task tk = tasks[find_task_wave(tasks, ntasks, gi)];
        const JobDesc &J = jobs[tk.job];
        if ((J.step == 3) != E8 || J.comb) continue;  // uniform per wave: another launch's job
        const int level = tk.ref;  // 1: sub-superblock table, 0: superblock table
        const int u = gi - tk.offset;
        const int X2 = J.xnb, W = J.W, H = J.H, QW = W + 2, step = J.step, nr = J.nrefs;
        const int XU = level ? J.xnb / 2 : J.xsb;
        const int xu = u % XU, yu = u / XU;
        const Blk b = block_limits(J.blen[level], J.bsep[level], J.boff[level], W, H, xu, yu);
        const int nq = level ? 4 : 16, side = level ? 2 : 4;
        int *out = level ? J.m1tab + (size_t)u * 24 : J.m0tab + (size_t)u * 32;
        if (b.empty()) {
            if (lane < (level ? 24 : 32)) out[lane] = 0;
            continue;
        }
        uint32_t *s_pic = s_mt + wv * ly.wave;
        uint32_t *s_slot = s_pic + ly.pd * ly.prows;
        // lane r*16 + q holds candidate q of reference r and the index of its
        // first occurrence among the reference's candidates
        const int cr = lane >> 4, cq = lane & 15;
        const bool cand = lane < 32 && cr < nr && cq < nq;
        dme_mv cv;
        cv.x = cv.y = 0;
        if (cand) cv = J.vec[cr][(size_t)(yu * side + cq / side) * X2 + xu * side + cq % side];
        int first = cq;
        for (int p = nq - 1; p >= 0; --p) {
            const int px = __shfl(cv.x, (lane & 48) + p, WAVE), py = __shfl(cv.y, (lane & 48) + p, WAVE);
            if (p < cq && px == cv.x && py == cv.y) first = p;
        }
        // picture samples: rows of pd dwords, zero beyond xl
        {
            const int ppad = (b.xl + 3) & ~3;
            const B8Rect pr = b8_rect(J.pic8, J.p8, s_pic, ly.pd, b.xp, b.yp, b.xl, b.yl, ppad);
            if (b8_inside(J.p8, H, b.xp, b.yp, b.yl, ppad)) {
                const int np = (b.yl + pr.rpp - 1) / pr.rpp;
                for (int p0 = 0; p0 < np; p0 += 4) {
                    uint32_t lo[4], hi[4];
                    b8_load(pr, J.p8, p0, lo, hi);
                    b8_store(pr, p0, lo, hi);
                }
            } else {
                b8_copy_clamped(pr, J.pic8, J.p8, W, H, b.xp, b.yp);
            }
        }
        // lane (row0 + k * rpp, col) of the unit in the SAD passes
        const int nd = (b.xl + 3) >> 2, rpp = WAVE / nd;
        const int row0 = (int)(((float)lane + 0.5f) * (1.0f / (float)nd)), col = lane - row0 * nd;
        const bool on = row0 < rpp;
        const int valid = b.xl - 4 * col;
        const uint32_t mask = valid >= 4 ? 0xffffffffu : ((1u << (8 * valid)) - 1u);
        const int npass = (b.yl + rpp - 1) / rpp;
        WAVE_SYNC();
        int mysum = 0;
        unsigned long long todo = __ballot(cand && first == cq);
        while (todo) {
            const int L = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            const int r = L >> 4, q = L & 15;
            dme_mv m;
            m.x = __builtin_amdgcn_readlane(cv.x, L);
            m.y = __builtin_amdgcn_readlane(cv.y, L);
            const UpStart us = up_start(b, m, step, 2 * W, 2 * H);
            const int xq = 2 * us.rsx + us.rx, yq = 2 * us.rsy + us.ry;  // quarter-pel start
            const int X0 = xq >> 2, Y0 = yq >> 2, fx = xq & 3, fy = yq & 3;
            const uint32_t sel = (uint32_t)fx | ((uint32_t)(fx + 4) << 8) | 0x0c0c0000u;
            const uint32_t *__restrict__ qp = J.qp[r];
            int xo[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) xo[i] = clampi(X0 + 4 * col + i, -1, W) + 1;
            uint32_t *slot = s_slot + (r * 4 + q) * ly.slot;
            uint32_t acc = 0;
            // eighth pel: each sample from four up samples (no single phase)
            const bool e8 = E8 && us.e;
            for (int p = 0; e8 && p < npass; ++p) {
                const int row = row0 + p * rpp;
                if (on && row < b.yl) {
                    uint32_t v = 0;
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        v |= (uint32_t)qp_eighth(qp, W, H, us.rsx + 2 * (4 * col + i), us.rsy + 2 * row, us.rx, us.ry)
                             << (8 * i);
                    v &= mask;
                    acc = __builtin_amdgcn_sad_u8(s_pic[row * ly.pd + col], v, acc);
                    if (level) slot[row * ly.spd + col] = v;
                }
            }
            for (int p0 = 0; !e8 && p0 < npass; p0 += 4) {
                uint32_t d[4][4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int row = row0 + (p0 + k) * rpp;
#pragma unroll
                    for (int i = 0; i < 4; ++i) d[k][i] = 0;
                    if (on && row < b.yl) {
                        const uint32_t *qr = qp + (size_t)((clampi(Y0 + row, -1, H) + 1) * 4 + fy) * QW;
#pragma unroll
                        for (int i = 0; i < 4; ++i) d[k][i] = qr[xo[i]];
                    }
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int row = row0 + (p0 + k) * rpp;
                    if (on && row < b.yl) {
                        const uint32_t lo = __builtin_amdgcn_perm(d[k][1], d[k][0], sel);
                        const uint32_t hi = __builtin_amdgcn_perm(d[k][3], d[k][2], sel);
                        const uint32_t v = __builtin_amdgcn_perm(hi, lo, 0x05040100u) & mask;
                        acc = __builtin_amdgcn_sad_u8(s_pic[row * ly.pd + col], v, acc);
                        if (level) slot[row * ly.spd + col] = v;
                    }
                }
            }
            const int sum = wave_sum((int)acc);
            if (lane == L) mysum = sum;
        }
        // every candidate takes the sum of its first occurrence
        const int fsum = __shfl(mysum, (lane & 48) + first, WAVE);
        if (cand) out[(level ? 4 : 16) * cr + cq] = fsum;
        if (level == 1 && nr > 1) {
            // bi-prediction SADs of the distinct pairs, then of every pair
            WAVE_SYNC();
            const unsigned long long d0 = __ballot(cand && cr == 0 && first == cq);
            const unsigned long long d1 = __ballot(cand && cr == 1 && first == cq) >> 16;
            for (unsigned long long m0 = d0; m0; m0 &= m0 - 1) {
                const int q1 = __ffsll((long long)m0) - 1;
                for (unsigned long long m1 = d1; m1; m1 &= m1 - 1) {
                    const int q2 = __ffsll((long long)m1) - 1;
                    const uint32_t *s0 = s_slot + q1 * ly.slot, *s1 = s_slot + (4 + q2) * ly.slot;
                    // BiBlockEighthPel reads a second reference whose window
                    // leaves the plane from the first one's start
                    dme_mv v1, v2m;
                    v1.x = __builtin_amdgcn_readlane(cv.x, q1);
                    v1.y = __builtin_amdgcn_readlane(cv.y, q1);
                    v2m.x = __builtin_amdgcn_readlane(cv.x, 16 + q2);
                    v2m.y = __builtin_amdgcn_readlane(cv.y, 16 + q2);
                    const UpStart u1 = up_start(b, v1, step, 2 * W, 2 * H);
                    const UpStart u2 = up_start(b, v2m, step, 2 * W, 2 * H);
                    const bool moved = E8 && u2.e && u2.bounds;
                    uint32_t acc = 0;
                    for (int p = 0; p < npass; ++p) {
                        const int row = row0 + p * rpp;
                        if (on && row < b.yl) {
                            const int o = row * ly.spd + col;
                            uint32_t v2 = s1[o];
                            if (moved) {
                                v2 = 0;
#pragma unroll
                                for (int i = 0; i < 4; ++i)
                                    v2 |= (uint32_t)qp_eighth(J.qp[1], W, H, u1.rsx + 2 * (4 * col + i), u1.rsy + 2 * row,
                                                              u2.rx, u2.ry)
                                          << (8 * i);
                                v2 &= mask;
                            }
                            acc = __builtin_amdgcn_sad_u8(s_pic[row * ly.pd + col], avg_up_u8(s0[o], v2), acc);
                        }
                    }
                    const int sum = wave_sum((int)acc);
                    if (lane == 0) s_pair[wv][q1 * 4 + q2] = sum;
                }
            }
            WAVE_SYNC();
            const int f0 = __shfl(first, (lane >> 2) & 3, WAVE), f1 = __shfl(first, 16 + (lane & 3), WAVE);
            if (lane < 16) out[8 + lane] = s_pair[wv][f0 * 4 + f1];
        }
    }
}

// ---------------------------------------------------------------------------
// Tables of the combined-component jobs, whose samples need 9 bits: the same
// records as k_subpel_bulk and k_mode_tables write, with every SAD summed
// over the picture's int16 samples and the int16 up plane (up16_at), one
// wave per block (sub-pel) or unit (mode decision).
// ---------------------------------------------------------------------------
// Sub-pel table: the block and the up-plane window every candidate reads
// are staged in LDS (clamped reads, as up16_at clamps), then lane c sums
// candidate c over the block.  LDS per wave: `words` int16 (host-sized for
// the largest block).
__host__ __device__ __forceinline__ int subpel16_window(int bl) { return 2 * bl + 4; }

__global__ __launch_bounds__(256) void k_subpel_table16(const JobDesc *__restrict__ jobs,
                                                        const Task *__restrict__ tasks, int ntasks, int total,
                                                        int words) {
    extern __shared__ int16_t s_t16[];
    const int wv = wave_id(), lane = threadIdx.x & 63;
    const int gi = xcd_block() * 4 + wv;
    if (gi >= total) return;  // each wave works alone
    const Task tk = tasks[find_task_wave(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    const int r = tk.ref, k = gi - tk.offset;
    const int X = J.xnb, W = J.W, H = J.H, UW = 2 * W, UH = 2 * H;
    const int xb = k % X, yb = k / X;
    const int prec = J.prec, rad = sub_radius(prec), Q = 2 * rad + 1, QS = Q * Q;
    const Blk b = block_limits(J.blen[2], J.bsep[2], J.boff[2], W, H, xb, yb);
    const dme_mv pel = J.vec[r][k];
    int *__restrict__ rec = J.qtab[r] + (size_t)wf_pos(xb, yb, X, J.ynb) * SUB_SLOT;
    if (!b.empty() && prec > 0) {
        // candidate c: m = (pel << tq) + (c % Q - rad, c / Q - rad) in units
        // of 1/2^tq pel (step tq), as k_subpel_bulk tabulates it; its up
        // samples lie in columns ux0 .. ux0 + UC - 1 and rows uy0 .. of the
        // up plane (the +1 neighbours of the quarter-pel weights included)
        const int tq = sub_unit(prec);
        const int mx0 = (pel.x << tq) - rad, my0 = (pel.y << tq) - rad;
        const int ux0 = 2 * b.xp + (tq == 2 ? (mx0 >> 1) : mx0), uy0 = 2 * b.yp + (tq == 2 ? (my0 >> 1) : my0);
        const int UC = subpel16_window(b.xl), UR = subpel16_window(b.yl);
        int16_t *__restrict__ s_pic = s_t16 + (size_t)wv * words;
        int16_t *__restrict__ s_up = s_pic + ((b.xl * b.yl + 1) & ~1);
        const int16_t *__restrict__ up = reinterpret_cast<const int16_t *>(J.qp[r]);
        const int16_t *__restrict__ pic = J.pic[0];
        for (int i = lane; i < UC * UR; i += WAVE) {
            const int row = i / UC, col = i - row * UC;
            s_up[i] = up[(size_t)clampi(uy0 + row, 0, UH - 1) * UW + clampi(ux0 + col, 0, UW - 1)];
        }
        for (int i = lane; i < b.xl * b.yl; i += WAVE) {
            const int row = i / b.xl, col = i - row * b.xl;
            s_pic[i] = pic[(size_t)(b.yp + row) * W + b.xp + col];
        }
        WAVE_SYNC();
        if (lane < QS) {
            dme_mv m;
            m.x = (pel.x << tq) + lane % Q - rad;
            m.y = (pel.y << tq) + lane / Q - rad;
            const UpStart us = up_start<false>(b, m, tq, UW, UH);
            const int rx = us.rx, ry = us.ry;
            const int w0 = (2 - rx) * (2 - ry), w1 = rx * (2 - ry), w2 = (2 - rx) * ry, w3 = rx * ry;
            const int16_t *q0 = s_up + (us.rsy - uy0) * UC + (us.rsx - ux0);
            int sad = 0;
            for (int dy = 0; dy < b.yl; ++dy) {
                const int16_t *q = q0 + 2 * dy * UC;
                const int16_t *pr = s_pic + dy * b.xl;
                for (int dx = 0; dx < b.xl; ++dx) {
                    const int v = w0 * q[2 * dx] + w1 * q[2 * dx + 1] + w2 * q[2 * dx + UC] + w3 * q[2 * dx + UC + 1];
                    sad += abs(((v + 2) >> 2) - (int)pr[dx]);
                }
            }
            rec[lane] = sad;
        }
    }
    if (lane == 0) {
        rec[49] = __float_as_int(J.lam[k]);
        rec[50] = pel.x;
        rec[51] = pel.y;
        const dme_cost pc = J.pc[r][k];
        rec[52] = __float_as_int(pc.SAD);
        rec[53] = __float_as_int(pc.mvcost);
        rec[54] = __float_as_int(pc.total);
    }
}

// SAD of one candidate of a combined-component job over a unit, lanes over
// its samples, 8 in flight: the candidate's weights are the wave's own, so
// the loads of zero-weight up samples are skipped (one load a sample at
// half pel, one to four at quarter and eighth pel).
template <bool E8>
__device__ int wave_up16_sad(const int16_t *__restrict__ pic, const int16_t *__restrict__ up, int W, int H,
                             const Blk &b, const UpStart &s) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int UW = 2 * W, UH = 2 * H;
    const bool e = E8 && s.e;
    const int one = e ? 4 : 2, rx = s.rx, ry = s.ry;
    const int w0 = (one - rx) * (one - ry), w1 = rx * (one - ry), w2 = (one - rx) * ry, w3 = rx * ry;
    const int n = b.xl * b.yl;
    int acc = 0;
    for (int b0 = 0; b0 < n; b0 += 8 * WAVE) {
        int p[8], v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int i = b0 + k * WAVE + lane;
            p[k] = v[k] = 0;
            if (i < n) {
                const int dy = i / b.xl, dx = i - dy * b.xl;
                const int uy = s.rsy + 2 * dy, ux = s.rsx + 2 * dx;
                const int y0 = clampi(uy, 0, UH - 1), x0 = clampi(ux, 0, UW - 1);
                const int y1 = clampi(uy + 1, 0, UH - 1), x1 = clampi(ux + 1, 0, UW - 1);
                int t = w0 * up[(size_t)y0 * UW + x0];
                if (rx) t += w1 * up[(size_t)y0 * UW + x1];
                if (ry) t += w2 * up[(size_t)y1 * UW + x0];
                if (rx && ry) t += w3 * up[(size_t)y1 * UW + x1];
                v[k] = e ? (t + 8) >> 4 : (t + 2) >> 2;
                p[k] = pic[(size_t)(b.yp + dy) * W + b.xp + dx];
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) acc += abs(v[k] - p[k]);
    }
    return wave_sum(acc);
}

template <bool E8>
__global__ __launch_bounds__(256) void k_mode_tables16(const JobDesc *__restrict__ jobs,
                                                       const Task *__restrict__ tasks, int ntasks, int total) {
    const int lane = threadIdx.x & 63;
    const int gi = xcd_block() * 4 + wave_id();
    if (gi >= total) return;  // each wave works alone
    const Task tk = tasks[find_task_wave(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    const int level = tk.ref;  // 1: sub-superblock table, 0: superblock table
    const int u = gi - tk.offset;
    const int X2 = J.xnb, W = J.W, H = J.H, step = J.step, nr = J.nrefs;
    const int XU = level ? J.xnb / 2 : J.xsb;
    const int xu = u % XU, yu = u / XU;
    const Blk b = block_limits(J.blen[level], J.bsep[level], J.boff[level], W, H, xu, yu);
    const int nq = level ? 4 : 16, side = level ? 2 : 4;
    int *out = level ? J.m1tab + (size_t)u * 24 : J.m0tab + (size_t)u * 32;
    if (b.empty()) {
        if (lane < (level ? 24 : 32)) out[lane] = 0;
        return;
    }
    // candidate q of reference r: the vector of block q of the unit
    // (candidates that no lane reads stay as they are, as k_mode_tables
    // leaves them)
    for (int r = 0; r < nr; ++r)
        for (int q = 0; q < nq; ++q) {
            const dme_mv m = J.vec[r][(size_t)(yu * side + q / side) * X2 + xu * side + q % side];
            const UpStart us = up_start<E8>(b, m, step, 2 * W, 2 * H);
            const int sd = wave_up16_sad<E8>(J.pic[0], reinterpret_cast<const int16_t *>(J.qp[r]), W, H, b, us);
            if (lane == 0) out[nq * r + q] = sd;
        }
    if (level == 1 && nr > 1) {
        // bi-prediction SADs of every pair (candidate q1 of reference 1,
        // candidate q2 of reference 2)
        for (int pr = 0; pr < 16; ++pr) {
            const int q1 = pr >> 2, q2 = pr & 3;
            const dme_mv v1 = J.vec[0][(size_t)(yu * 2 + q1 / 2) * X2 + xu * 2 + q1 % 2];
            const dme_mv v2 = J.vec[1][(size_t)(yu * 2 + q2 / 2) * X2 + xu * 2 + q2 % 2];
            const int s = wave_bi_sad<E8, true>(J, b, v1, v2);
            if (lane == 0) out[8 + pr] = s;
        }
    }
}

// ModeDecider::DoME for one unit and reference (me_mode_decn.cpp:279-364;
// BlockMatcher::FindBestMatchSubp, block_match.cpp:277-325) from tabulated
// SADs, one (unit, reference) per lane.  g are the four guide vectors with
// their SADs; duplicates are skipped as AddNewVlist does (they cannot win a
// strict-< comparison anyway).  Totals that need the per-pixel float chain
// are computed by the whole wave, so every lane of the wave calls this.
// Returns the index of the chosen guide (-1 when nothing was chosen).
template <bool E8, bool COMB>
__device__ int unit_me(const JobDesc &J, bool on, const Blk &b, int r, dme_mv pred, const dme_mv g[4],
                       const int gs[4], float lambda, dme_mv &out_mv, dme_cost &out_cost) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int UW = 2 * J.W, UH = 2 * J.H;
    float tot[4], mvc[4], st[4];
    bool cand[4], need[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        bool dup = false;
#pragma unroll
        for (int p = 0; p < q; ++p) dup = dup || (g[p].x == g[q].x && g[p].y == g[q].y);
        cand[q] = on && !b.empty() && !dup;
        mvc[q] = var_up(pred, g[q]);
        st[q] = mvc[q] * lambda;
        tot[q] = (float)gs[q] + st[q];
        const UpStart us = up_start<E8>(b, g[q], J.step, UW, UH);
        need[q] = cand[q] && (us.bounds || us.e || J.fp_sem == DME_FP_C) && !chain_is_exact(st[q], gs[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        unsigned long long mm = __ballot(need[q]);
        while (mm) {
            const int src = __ffsll((long long)mm) - 1;
            mm &= mm - 1;
            Blk bs;
            bs.xp = __builtin_amdgcn_readlane(b.xp, src);
            bs.yp = __builtin_amdgcn_readlane(b.yp, src);
            bs.xl = __builtin_amdgcn_readlane(b.xl, src);
            bs.yl = __builtin_amdgcn_readlane(b.yl, src);
            dme_mv m;
            m.x = __builtin_amdgcn_readlane(g[q].x, src);
            m.y = __builtin_amdgcn_readlane(g[q].y, src);
            const float start = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(st[q]), src));
            const int rs = __builtin_amdgcn_readlane(r, src);
            const UpStart us = up_start<E8>(bs, m, J.step, UW, UH);
            const float v = wave_chain<E8, COMB>(start, J.pic[0], J.W, rs ? J.qp[1] : J.qp[0], UW, UH, bs, us);
            if (lane == src) tot[q] = v;
        }
    }
    dme_cost best;
    best.SAD = 0.0f;
    best.mvcost = 0.0f;
    best.total = 100000000.0f;
    dme_mv bmv;
    bmv.x = 0;
    bmv.y = 0;
    int bq = -1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (cand[q] && tot[q] < best.total) {
            bmv = g[q];
            best.total = tot[q];
            best.mvcost = mvc[q];
            best.SAD = tot[q] - st[q];
            bq = q;
        }
    }
    out_mv = bmv;
    out_cost = best;
    return bq;
}

// DoUnitDecn for levels 1 and 0 (me_mode_decn.cpp:368-441; no intra).
__device__ float unit_decn(const JobDesc &J, int level, float loc_lambda, dme_cost c[2], int bisad, int &mode) {
    const float lf = J.lf[level];
    const float mode_cost = 0.0f;
    mode = DME_REF1_ONLY;
    c[0].total *= lf;
    float min_cost = c[0].total + mode_cost;
    if (J.nrefs > 1) {
        c[1].total *= lf;
        float uc = c[1].total + mode_cost;
        if (uc < min_cost) {
            mode = DME_REF2_ONLY;
            min_cost = uc;
        }
        dme_cost bc;
        bc.mvcost = c[0].mvcost + c[1].mvcost;
        bc.SAD = (float)bisad;
        bc.total = bc.SAD + loc_lambda * bc.mvcost;
        bc.total *= lf;
        uc = bc.total + mode_cost;
        if (uc < min_cost) {
            mode = DME_REF1AND2;
            min_cost = uc;
        }
    }
    return min_cost;
}

// ---------------------------------------------------------------------------
// Superblock decisions (ModeDecider::DoSBDecn / DoLevelDecn,
// me_mode_decn.cpp:161-276).  One workgroup per picture, wavefront over
// superblocks t = X + Y, one wave per superblock.  Level 0 always runs: the
// superblock cost after level 1 never exceeds the level-2 cost.  Everything a
// superblock reads from memory is fetched in one round at the start (the
// vectors its predictions use belong to superblocks of earlier steps, or are
// its own level-2 vectors, which it overwrites only after reading them);
// values then move between lanes by shuffles.
// ---------------------------------------------------------------------------
__device__ __forceinline__ dme_mv shfl_mv(dme_mv v, int src) {
    dme_mv o;
    o.x = __shfl(v.x, src, WAVE);
    o.y = __shfl(v.y, src, WAVE);
    return o;
}
__device__ __forceinline__ dme_cost shfl_cost(dme_cost c, int src) {
    dme_cost o;
    o.SAD = __shfl(c.SAD, src, WAVE);
    o.mvcost = __shfl(c.mvcost, src, WAVE);
    o.total = __shfl(c.total, src, WAVE);
    return o;
}
__device__ __forceinline__ float readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

#ifndef MODE_NB
#define MODE_NB 4
#endif
// E8: the eighth-pel jobs of the launch (the others exit), else the rest;
// COMB likewise for the combined-component jobs.
template <bool E8, bool COMB>
__global__ __launch_bounds__(1024) void k_mode_wave(const JobDesc *__restrict__ jobs) {
    const JobDesc &J = jobs[blockIdx.x];
    if ((J.step == 3) != E8 || (J.comb != 0) != COMB) return;  // uniform per workgroup
    const int XS = J.xsb, YS = J.ysb, X2 = J.xnb, X1 = J.xnb / 2;
    const int wave = wave_id(), nwaves = blockDim.x >> 6, lane = threadIdx.x & 63;
    const int nr = J.nrefs;
    const dme_mv *__restrict__ V0 = J.vec[0];
    const dme_mv *__restrict__ V1 = J.vec[1];
    for (int t = 0; t <= (XS - 1) + (YS - 1); ++t) {
        const int ylo = t - (XS - 1) > 0 ? t - (XS - 1) : 0;
        const int yhi = min(YS - 1, t);
        for (int ys = ylo + wave; ys <= yhi; ys += nwaves) {
            const int xs = t - ys;
            const size_t s = (size_t)ys * XS + xs;
            // ---- loads.  lanes 0..15: level-2 costs of the 16 blocks;
            // lanes 0..7: unit u = lane/2 against reference r = lane%2 at level 1;
            // every lane: one level-1 bi-prediction SAD (unit lane/16, pair lane%16);
            // lanes 0..31: level-0 candidate SADs; lanes 0..1: level-0 prediction
            float ucl = 0.0f;
            if (lane < 16) ucl = J.ucost[(size_t)(4 * ys + (lane >> 2)) * X2 + 4 * xs + (lane & 3)];
            const int u = (lane >> 1) & 3, r1 = lane & 1;
            const int xu = 2 * xs + (u & 1), yu = 2 * ys + (u >> 1);
            const size_t ku = (size_t)yu * X1 + xu;
            const bool on1 = lane < 8 && r1 < nr;
            dme_mv g[4];
            int gs[4];
            float lam1 = 0.0f;
            dme_mv pred1;
            pred1.x = pred1.y = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                g[q].x = g[q].y = 0;
                gs[q] = 0;
            }
            if (on1) {
                const dme_mv *__restrict__ V = r1 ? V1 : V0;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    g[q] = V[(size_t)(2 * yu + (q >> 1)) * X2 + 2 * xu + (q & 1)];
                    gs[q] = J.m1tab[ku * 24 + r1 * 4 + q];
                }
                lam1 = J.lam1[ku];
                pred1 = mode_pred(V, X2, 2 * xu, 2 * yu);
            }
            int bitab = 0;
            if (nr > 1) {
                const int ub = lane >> 4;
                const size_t kb = (size_t)(2 * ys + (ub >> 1)) * X1 + 2 * xs + (ub & 1);
                bitab = J.m1tab[kb * 24 + 8 + (lane & 15)];
            }
            const int m0v = lane < 32 ? J.m0tab[s * 32 + lane] : 0;
            const float lam0 = J.lam0[s];
            dme_mv pred0;
            pred0.x = pred0.y = 0;
            if (lane < nr) pred0 = mode_pred(lane ? V1 : V0, X2, 4 * xs, 4 * ys);

            // ---- level 2: sum of the block costs in raster order
            float sbc = 0.0f;
#pragma unroll
            for (int i = 0; i < 16; ++i) sbc += readlane_f(ucl, i);
            int split = 2;

            // ---- level 1: the eight (unit, reference) searches
            const Blk b1 = block_limits(J.blen[1], J.bsep[1], J.boff[1], J.W, J.H, xu, yu);
            dme_mv lv;
            dme_cost lc;
            const int lq = unit_me<E8, COMB>(J, on1, b1, r1, pred1, g, gs, lam1, lv, lc);
            // unit decisions, lane u for unit u
            const int sa = (2 * lane) & 63, sb = (2 * lane + 1) & 63;
            dme_cost cu[2] = {shfl_cost(lc, sa), shfl_cost(lc, sb)};
            const int qa = __shfl(lq, sa, WAVE), qb = __shfl(lq, sb, WAVE);
            const float lamu = __shfl(lam1, sa, WAVE);
            const int bsrc = ((lane & 3) * 16 + (qa > 0 ? qa : 0) * 4 + (qb > 0 ? qb : 0)) & 63;
            const int bv = __shfl(bitab, bsrc, WAVE);
            const int bis = (nr > 1 && qa >= 0 && qb >= 0) ? bv : 0;
            int mu;
            const float ucost1 = unit_decn(J, 1, lamu, cu, bis, mu);
            float sb1 = 0.0f;
#pragma unroll
            for (int i = 0; i < 4; ++i) sb1 += readlane_f(ucost1, i);
            {
                // block lane of the superblock: its unit and that unit's result
                const int bu = ((lane >> 3) << 1) | ((lane >> 1) & 1);
                const int mb = __shfl(mu, bu & 3, WAVE);
                const dme_mv va = shfl_mv(lv, (2 * bu) & 63), vb = shfl_mv(lv, (2 * bu + 1) & 63);
                if (sb1 <= sbc) {
                    sbc = sb1;
                    split = 1;
                    if (lane < 16) {
                        const size_t d = (size_t)(4 * ys + (lane >> 2)) * X2 + 4 * xs + (lane & 3);
                        J.mode[d] = mb;
                        J.dc[0][d] = 0;
                        J.vec[0][d] = va;
                        if (nr > 1) J.vec[1][d] = vb;
                    }
                }
            }

            // ---- level 0: guides are the level-1 results; their SADs are
            // those of the blocks they came from (a unit with no result offers
            // (0,0), whose SAD the wave computes)
            const Blk b0 = block_limits(J.blen[0], J.bsep[0], J.boff[0], J.W, J.H, xs, ys);
            const int r0 = lane & 1;
            dme_mv g0[4];
            int gs0[4];
            bool miss[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                g0[q] = shfl_mv(lv, (2 * q + r0) & 63);
                const int bq = __shfl(lq, (2 * q + r0) & 63, WAVE);
                const int e = r0 * 16 + (2 * (q >> 1) + (bq > 0 ? bq >> 1 : 0)) * 4 + 2 * (q & 1) + (bq > 0 ? bq & 1 : 0);
                const int tv = __shfl(m0v, e & 63, WAVE);
                gs0[q] = bq >= 0 ? tv : 0;
                miss[q] = lane < nr && bq < 0 && !b0.empty();
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                unsigned long long mm = __ballot(miss[q]);
                while (mm) {
                    const int src = __ffsll((long long)mm) - 1;
                    mm &= mm - 1;
                    dme_mv m;
                    m.x = __builtin_amdgcn_readlane(g0[q].x, src);
                    m.y = __builtin_amdgcn_readlane(g0[q].y, src);
                    int sd = 0;
                    wave_up_sads<E8, COMB>(J.pic[0], J.W, src ? J.qp[1] : J.qp[0], 2 * J.W, 2 * J.H, b0, J.step, &m, 1,
                                           &sd);
                    if (lane == src) gs0[q] = sd;
                }
            }
            dme_mv mv0;
            dme_cost mc0;
            unit_me<E8, COMB>(J, lane < nr, b0, r0, pred0, g0, gs0, lam0, mv0, mc0);
            const dme_mv va0 = shfl_mv(mv0, 0), vb0 = shfl_mv(mv0, 1);
            dme_cost c0[2] = {shfl_cost(mc0, 0), shfl_cost(mc0, 1)};
            int bis0 = 0;
            if (nr > 1 && !b0.empty())
                bis0 = E8 || COMB ? wave_bi_sad<E8, COMB, MODE_NB>(J, b0, va0, vb0) : wave_bi_sad_u8(J, b0, va0, vb0);
            int m0;
            const float sb0 = unit_decn(J, 0, lam0, c0, bis0, m0);
            if (sb0 <= sbc) {
                sbc = sb0;
                split = 0;
                if (lane < 16) {
                    const size_t d = (size_t)(4 * ys + (lane >> 2)) * X2 + 4 * xs + (lane & 3);
                    J.mode[d] = m0;
                    J.dc[0][d] = 0;
                    J.vec[0][d] = va0;
                    if (nr > 1) J.vec[1][d] = vb0;
                }
            }
            if (lane == 0) {
                J.sbc[s] = sbc;
                J.split[s] = split;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Learned proposal stage (no counterpart in the reference).  The plain pel
// search runs down to pyramid level 2 (quarter resolution); a small CNN on
// the level-0 block grid (dme_learned.hip) turns the level-2 vectors, their
// costs and the picture's texture into four candidate vectors per block and
// reference; exact integer SADs over a +-2 pel window around each candidate
// pick the level-0 pel vector, with a wider exact search where the result
// costs more than 1.2 times the network's predicted SAD.  The vectors then go
// through the unchanged sub-pel and mode-decision stages.
// ---------------------------------------------------------------------------
#define LP_SCALE_PEL 8.0f  // pel per unit of the network's offset outputs

__device__ __forceinline__ uint16_t f2bf_rne(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// level-2 vector and SAD covering level-0 block (x, y)
__device__ __forceinline__ void lp_guide(const JobDesc &J, int r, int x, int y, dme_mv &v2, float &s2) {
    const int X2 = J.lxb[2], Y2 = J.lyb[2];
    const size_t k2 = (size_t)min(y >> 2, Y2 - 1) * X2 + min(x >> 2, X2 - 1);
    v2 = J.lvec[r][2][k2];
    s2 = J.lpc[r][2][k2].SAD;
}

// Input features of the network, one position per level-0 block:
//   3r + 0, 1: level-2 vector in pel / 32; 3r + 2: log(1 + level-2 SAD) / 8;
//   6, 7: mean absolute horizontal / vertical gradient of the level-2
//   picture over the 4 x 4 samples under the block, / 32; the rest zero.
// With labels: target offset of the plain level-0 vector from 4 x the
// level-2 vector, log(1 + its SAD), valid.
__global__ __launch_bounds__(256) void k_lp_features(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                     int ntasks, int total, uint16_t *__restrict__ feat,
                                                     float *__restrict__ lab) {
    const int gi = blockIdx.x * 256 + threadIdx.x;
    if (gi >= total) return;
    const Task tk = tasks[find_task(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    const int k = gi - tk.offset, X = J.xnb, x = k % X, y = k / X;
    const size_t p = (size_t)J.lp_off + k;
    float f[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) f[c] = 0.0f;
    dme_mv v2[2];
    v2[0].x = v2[0].y = v2[1].x = v2[1].y = 0;
    for (int r = 0; r < J.nrefs; ++r) {
        float s2;
        lp_guide(J, r, x, y, v2[r], s2);
        f[3 * r] = 4.0f * v2[r].x / 32.0f;
        f[3 * r + 1] = 4.0f * v2[r].y / 32.0f;
        f[3 * r + 2] = log1pf(s2) / 8.0f;
    }
    {
        const int W2 = J.lw[2], H2 = J.lh[2];
        const int cx = (x * J.bsep[2][0] + J.bsep[2][0] / 2) >> 2, cy = (y * J.bsep[2][1] + J.bsep[2][1] / 2) >> 2;
        const int16_t *__restrict__ pl = J.pic[2];
        int gx = 0, gy = 0;
        for (int dy = -2; dy < 2; ++dy)
            for (int dx = -2; dx < 2; ++dx) {
                const int yy = clampi(cy + dy, 0, H2 - 1), xx = clampi(cx + dx, 0, W2 - 1);
                const int v = pl[(size_t)yy * W2 + xx];
                gx += abs(pl[(size_t)yy * W2 + clampi(xx + 1, 0, W2 - 1)] - v);
                gy += abs(pl[(size_t)clampi(yy + 1, 0, H2 - 1) * W2 + xx] - v);
            }
        f[6] = gx / (16.0f * 32.0f);
        f[7] = gy / (16.0f * 32.0f);
    }
    uint16_t *fo = feat + p * 32;
#pragma unroll
    for (int c = 0; c < 32; ++c) fo[c] = c < 8 ? f2bf_rne(f[c]) : (uint16_t)0;
    if (lab) {
        const Blk b = block_limits(J.blen[2], J.bsep[2], J.boff[2], J.W, J.H, x, y);
        float *lo = lab + p * 8;
        for (int r = 0; r < 2; ++r) {
            const bool valid = r < J.nrefs && !b.empty();
            const dme_mv v0 = valid ? J.vec[r][k] : v2[r];
            lo[4 * r] = (float)(v0.x - 4 * v2[r].x);
            lo[4 * r + 1] = (float)(v0.y - 4 * v2[r].y);
            lo[4 * r + 2] = valid ? log1pf(J.pc[r][k].SAD) : 0.0f;
            lo[4 * r + 3] = valid ? 1.0f : 0.0f;
        }
    }
}

// Exact refinement of the proposals.  Per (block, reference) item: the four
// hypotheses are clamped to within LP_BOX - 2 pel of 4 x the level-2 vector
// g; the candidates are the union of the +-2 windows around them; the winner
// is the smallest key SAD * 1024 + raster index in the box g +- LP_BOX.  If
// its SAD exceeds 1.2 x the predicted SAD, the +-4 window around g and the
// +-2 window around zero (box raster index, or 625 + window index when zero's
// window leaves the box) join in and replace the winner only with a smaller
// SAD.  Keys are distinct per position, so a position evaluated twice (in two
// overlapping windows) leaves the minimum unchanged.
//
// A wave takes LP_ITEMS consecutive items.  The picture block of each is
// staged from the byte plane (realigned, zero past its width), its box as
// whole dwords from the dword at or before the box's first column (a box
// that leaves the plane's margin is copied sample by sample with clamped
// coordinates).  Twenty lanes serve an item, lane (h, c) taking column c - 2
// of hypothesis h's window: it slides down the window's rows, realigns each
// box row it reads once and adds it to the SADs of the five candidates of its
// column whose block rows it meets (lp_slide).  The fallback windows are
// served the same way, a lane taking five candidate rows of one column.
// Writes the level-0 vector and SAD; the slot's counters (kLpStatCopies copies
// of {blocks, fallbacks}) count them.
#define LP_BOX 12
#define LP_BD (2 * LP_BOX + 1)
#define LP_ITEMS 3
#define LP_LANES 20
#define LP_FB_TASKS 23  // fallback lanes per item: 9 columns x 2 row bands of g's window, 5 columns of zero's

// LDS of one wave of k_lp_refine (dwords): per item the picture block (prows
// rows of pp dwords, rows 16-byte aligned), the box (brows rows of rp) and
// zero's window (zrows rows of zp); then the items' key minima (8 ints).
struct LpLds {
    int pp, prows, rp, brows, zp, zrows;
    __device__ int item() const { return (pp * prows + rp * brows + zp * zrows + 3) & ~3; }
    __device__ int wave() const { return LP_ITEMS * item() + 8; }
};

// Index of offset (dx, dy) in win_off's order for radius 2.
__device__ __forceinline__ int win_index2(int dx, int dy) {
    if (dy == 0) return dx == 0 ? 0 : (dx > 0 ? 2 * dx - 1 : -2 * dx);
    return 5 + (abs(dy) - 1) * 10 + (dx + 2) * 2 + (dy < 0 ? 1 : 0);
}

// SADs of five vertically adjacent candidates against the block whose packed
// rows are at pic (pitch pp dwords, bytes past the block's width zero):
// candidate dy (0..4) pairs block row y with source row row0 + dy + y, whose
// bytes start at byte offset bo (which may include a staging shift).  Each
// source row is read and realigned once and added to every candidate whose
// block row it meets.  NR, ND: rows and dwords of a whole block known at
// compile time; NR = 0 takes yl rows and nd dwords at run time, the last
// realigned dword masked by lmask.
template <int NR, int ND>
__device__ __forceinline__ void lp_slide(const uint32_t *__restrict__ pic, int pp, const uint32_t *__restrict__ src,
                                         int sp, int bo, int row0, int yl, int nd, uint32_t lmask, uint32_t acc[5]) {
    const uint32_t *rb = src + row0 * sp + (bo >> 2);
    const unsigned sh = bo & 3;
    if constexpr (NR > 0) {
        // rows stay in order: each row's loads are issued one iteration
        // ahead and no further (left alone, the compiler issues every row's
        // loads first and exhausts the registers)
        uint32_t ring[5][ND], r[ND + 1];
#pragma unroll
        for (int d = 0; d <= ND; ++d) r[d] = rb[d];
#pragma unroll
        for (int j = 0; j < NR + 4; ++j) {
            __builtin_amdgcn_sched_barrier(0);
            if (j < NR) {
                // rows are 16-byte aligned and at least four dwords long
                const uint4 q = *reinterpret_cast<const uint4 *>(pic + j * pp);
                const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                for (int d = 0; d < (ND < 4 ? ND : 4); ++d) ring[j % 5][d] = w[d];
                if constexpr (ND > 4) {
                    const uint2 q2 = *reinterpret_cast<const uint2 *>(pic + j * pp + 4);
                    ring[j % 5][4] = q2.x;
                    ring[j % 5][5] = q2.y;
                }
            }
            uint32_t a[ND];
#pragma unroll
            for (int d = 0; d < ND; ++d) a[d] = __builtin_amdgcn_alignbyte(r[d + 1], r[d], sh);
            if (j + 1 < NR + 4) {
#pragma unroll
                for (int d = 0; d <= ND; ++d) r[d] = rb[(j + 1) * sp + d];
            }
#pragma unroll
            for (int dy = 0; dy < 5; ++dy) {
                const int y = j - dy;
                if (y >= 0 && y < NR) {
#pragma unroll
                    for (int d = 0; d < ND; ++d) acc[dy] = __builtin_amdgcn_sad_u8(ring[y % 5][d], a[d], acc[dy]);
                }
            }
            // the row's sums are formed here, not deferred past later rows' loads
            asm volatile("" : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(acc[4]));
        }
    } else {
        for (int j = 0; j < yl + 4; ++j) {
            uint32_t r[9], a[8];
#pragma unroll
            for (int d = 0; d < 9; ++d) r[d] = d <= nd ? rb[j * sp + d] : 0u;
#pragma unroll
            for (int d = 0; d < 8; ++d) {
                a[d] = __builtin_amdgcn_alignbyte(r[d + 1], r[d], sh);
                if (d == nd - 1) a[d] &= lmask;
            }
#pragma unroll
            for (int dy = 0; dy < 5; ++dy) {
                const int y = j - dy;
                if (y >= 0 && y < yl) {
                    const uint32_t *pr = pic + y * pp;
#pragma unroll
                    for (int d = 0; d < 8; ++d)
                        if (d < nd) acc[dy] = __builtin_amdgcn_sad_u8(pr[d], a[d], acc[dy]);
                }
            }
        }
    }
}

// Smallest key of a lane's five candidates (dy < ndy count): index idx0 +
// dy * LP_BD in the box, or, for zero's window staged apart (zx >= -2), 625 +
// its window index of (zx, dy - 2).
template <int NR, int ND>
__device__ __forceinline__ int lp_column_key(const uint32_t *pic, int pp, const uint32_t *src, int sp, int bo, int row0,
                                             int yl, int nd, uint32_t lmask, int ndy, int idx0, int zx) {
    uint32_t acc[5] = {0u, 0u, 0u, 0u, 0u};
    lp_slide<NR, ND>(pic, pp, src, sp, bo, row0, yl, nd, lmask, acc);
    int key = INT_MAX;
#pragma unroll
    for (int dy = 0; dy < 5; ++dy) {
        const int idx = zx >= -2 ? LP_BD * LP_BD + win_index2(zx, dy - 2) : idx0 + dy * LP_BD;
        if (dy < ndy) key = min(key, (int)acc[dy] * 1024 + idx);
    }
    return key;
}

__global__ __launch_bounds__(256) void k_lp_refine(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                   int ntasks, int total, const float *__restrict__ out4,
                                                   unsigned long long *__restrict__ stats, LpLds ly) {
    extern __shared__ uint32_t s_lp[];
    const int wv = wave_id(), lane = threadIdx.x & (WAVE - 1);
    const int gbase = (xcd_block() * 4 + wv) * LP_ITEMS;
    if (gbase >= total) return;
    const int idw = ly.item();
    uint32_t *s_w = s_lp + wv * ly.wave();
    int *s_key = reinterpret_cast<int *>(s_w + LP_ITEMS * idw);  // [i] main minimum, [4 + i] fallback minimum
    if (lane < 8) s_key[lane] = INT_MAX;
    const int box_off = ly.pp * ly.prows, zero_off = box_off + ly.rp * ly.brows;

    // The lanes of group i = min(lane / 20, 2) hold the descriptor of item
    // gbase + i: job, reference, block index and limits, 4 x the level-2
    // vector, the first network output of the block, the predicted SAD, the
    // planes, and whether the box is copied as whole dwords (else sample by
    // sample, clamped) with its staging shift.
    const int me = min(lane / LP_LANES, LP_ITEMS - 1);
    int d_job, d_r, d_k, d_xp, d_yp, d_xl, d_yl, d_gx, d_gy, d_o, d_fast, d_sh, d_on, d_p8, d_W, d_H;
    const uint8_t *d_pic8, *d_ref8;
    float d_pred;
    {
        int c0 = 0, c1 = 0, c2 = 0;  // tasks starting at or before each item
        for (int j0 = 0; j0 < ntasks; j0 += WAVE) {
            const int j = j0 + lane;
            const int off = j < ntasks ? tasks[j].offset : INT_MAX;
            c0 += __popcll(__ballot(off <= gbase));
            c1 += __popcll(__ballot(off <= min(gbase + 1, total - 1)));
            c2 += __popcll(__ballot(off <= min(gbase + 2, total - 1)));
        }
        const int gi = min(gbase + me, total - 1);
        const Task tk = tasks[(me == 0 ? c0 : (me == 1 ? c1 : c2)) - 1];
        const JobDesc &J = jobs[tk.job];
        const int r = tk.ref, k = gi - tk.offset, X = J.xnb, x = k % X, y = k / X;
        const Blk b = block_limits(J.blen[2], J.bsep[2], J.boff[2], J.W, J.H, x, y);
        dme_mv g;
        float s2;
        lp_guide(J, r, x, y, g, s2);
        d_job = tk.job;
        d_r = r;
        d_k = k;
        d_xp = b.xp;
        d_yp = b.yp;
        d_xl = b.xl;
        d_yl = b.yl;
        d_gx = 4 * g.x;
        d_gy = 4 * g.y;
        d_o = (J.lp_off + k) * 32;
        d_on = gbase + me < total && !b.empty();
        d_pred = expm1f(fminf(out4[(size_t)d_o + 16 + r], 20.0f));
        d_p8 = J.p8;
        d_W = J.W;
        d_H = J.H;
        d_pic8 = J.pic8;
        d_ref8 = J.ref8[r];
        const int bx0 = b.xp + d_gx - LP_BOX, by0 = b.yp + d_gy - LP_BOX, RH = b.yl + 2 * LP_BOX;
        const int ax0 = bx0 & ~3, ncol = (b.xl + 2 * LP_BOX + (bx0 & 3) + 3) >> 2;
        d_fast = ax0 >= -B8_MG && ax0 + 4 * ncol <= J.p8 - B8_MG && by0 >= -B8_MG && by0 + RH <= J.H + B8_MG;
        d_sh = d_fast ? bx0 & 3 : 0;
    }
    // uniform copy of item i's value
    auto U = [&](int v, int i) { return __builtin_amdgcn_readlane(v, i * LP_LANES); };
    const unsigned long long onm = __ballot(lane % LP_LANES == 0 && d_on);
    const int onbits = (int)((onm & 1ull) | (((onm >> 20) & 1ull) << 1) | (((onm >> 40) & 1ull) << 2));
    // every active item a whole 16 x 16 (24 x 24) block: the slide's
    // unrolled forms
    const bool full16 = __ballot(d_on && (d_xl != 16 || d_yl != 16)) == 0;
    const bool full24 = !full16 && __ballot(d_on && (d_xl != 24 || d_yl != 24)) == 0;

    // staging: the 20 lanes of each item copy its picture block (realigned
    // from the byte plane, zero past xl) and its box (whole dwords from the
    // dword at or before its first column, or sample by sample with clamped
    // coordinates), all items at once
    if (lane < LP_ITEMS * LP_LANES && d_on) {
        const int q = lane - me * LP_LANES, bp = d_p8, nd = (d_xl + 3) >> 2;
        uint32_t *s_pic = s_w + me * idw, *s_box = s_pic + box_off;
        {
            const int n = d_yl * nd, psh = d_xp & 3;
            const float inv = 1.0f / (float)nd;
            const bool inside = b8_inside(bp, d_H, d_xp, d_yp, d_yl, 4 * nd);
            const uint8_t *pb = d_pic8 + (ptrdiff_t)d_yp * bp + (d_xp - psh);
            for (int e0 = 0; e0 < n; e0 += 4 * LP_LANES) {
                uint32_t lo[4], hi[4];
                int at[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int e = min(e0 + q + u * LP_LANES, n - 1);  // the last element repeats: no guard
                    const int row = (int)(((float)e + 0.5f) * inv), col = e - row * nd;
                    at[u] = row * ly.pp + col;
                    if (inside) {
                        const uint32_t *w = reinterpret_cast<const uint32_t *>(pb + (ptrdiff_t)row * bp) + col;
                        lo[u] = w[0];
                        hi[u] = w[1];
                    } else {
                        const uint8_t *src = d_pic8 + (ptrdiff_t)clampi(d_yp + row, 0, d_H - 1) * bp;
                        uint32_t v = 0;
#pragma unroll
                        for (int kk = 0; kk < 4; ++kk) v |= (uint32_t)src[clampi(d_xp + 4 * col + kk, 0, d_W - 1)] << (8 * kk);
                        lo[u] = v;
                        hi[u] = 0;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int e = min(e0 + q + u * LP_LANES, n - 1);
                    const int row = (int)(((float)e + 0.5f) * inv), col = e - row * nd;
                    const int valid = d_xl - 4 * col;
                    const uint32_t mask = valid >= 4 ? 0xffffffffu : ((1u << (8 * valid)) - 1u);
                    const uint32_t v = inside ? __builtin_amdgcn_alignbyte(hi[u], lo[u], psh) : lo[u];
                    s_pic[at[u]] = v & mask;
                }
            }
        }
        {
            const int bx0 = d_xp + d_gx - LP_BOX, by0 = d_yp + d_gy - LP_BOX;
            const int RW = d_xl + 2 * LP_BOX, RH = d_yl + 2 * LP_BOX;
            if (d_fast) {
                const int ncol = (RW + d_sh + 3) >> 2, n = RH * ncol;
                const float inv = 1.0f / (float)ncol;
                const uint32_t *__restrict__ src =
                    reinterpret_cast<const uint32_t *>(d_ref8 + (ptrdiff_t)by0 * bp + (bx0 - d_sh));
                const int dp = bp >> 2;
                for (int e0 = 0; e0 < n; e0 += 12 * LP_LANES) {
                    uint32_t v[12];
#pragma unroll
                    for (int u = 0; u < 12; ++u) {
                        const int e = min(e0 + q + u * LP_LANES, n - 1);  // the last element repeats: no guard
                        const int row = (int)(((float)e + 0.5f) * inv), col = e - row * ncol;
                        v[u] = src[(ptrdiff_t)row * dp + col];
                    }
#pragma unroll
                    for (int u = 0; u < 12; ++u) {
                        const int e = min(e0 + q + u * LP_LANES, n - 1);
                        const int row = (int)(((float)e + 0.5f) * inv), col = e - row * ncol;
                        s_box[row * ly.rp + col] = v[u];
                    }
                }
            } else {
                const int ncol = (RW + 3) >> 2, n = RH * ncol;
                const float inv = 1.0f / (float)ncol;
                for (int e = q; e < n; e += LP_LANES) {
                    const int row = (int)(((float)e + 0.5f) * inv), col = e - row * ncol;
                    const uint8_t *src = d_ref8 + (ptrdiff_t)clampi(by0 + row, 0, d_H - 1) * bp;
                    uint32_t v = 0;
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk)
                        if (4 * col + kk < RW) v |= (uint32_t)src[clampi(bx0 + 4 * col + kk, 0, d_W - 1)] << (8 * kk);
                    s_box[row * ly.rp + col] = v;
                }
            }
        }
    }
    WAVE_SYNC();

    // One loop serves the main windows (pass 0: lane = item x 20 + 5h + c)
    // and, when an item falls back, its fallback windows (pass 1: task t =
    // fallback item x 23 + q over the lanes), so that each form of lp_slide
    // is instantiated once.
    int best_sad[LP_ITEMS], best_idx[LP_ITEMS], nfb = 0, fbm = 0;
    for (int pass = 0, t0 = 0;;) {
        const int t = t0 + lane;
        int it = LP_ITEMS, src_off = box_off, sp = ly.rp, bo = 0, row0 = 0, ndy = 5, idx0 = 0, zx = -99, slot = 0;
        if (pass == 0) {
            it = lane / LP_LANES;
        } else if (t < nfb * LP_FB_TASKS) {
            const int f = t / LP_FB_TASKS;
            const int m1 = fbm & (fbm - 1), m2 = m1 & (m1 - 1);  // the f-th fallback item
            it = __builtin_ctz(f == 0 ? fbm : (f == 1 ? m1 : m2));
        }
        // the item's fields, by lane shuffles with every lane active (a
        // shuffle from an inactive lane does not return its value)
        const int src_lane = min(it, LP_ITEMS - 1) * LP_LANES;
        const int gx = __shfl(d_gx, src_lane), gy = __shfl(d_gy, src_lane), sh = __shfl(d_sh, src_lane);
        const int xl = __shfl(d_xl, src_lane), yl = __shfl(d_yl, src_lane), on = __shfl(d_on, src_lane);
        const int obase = __shfl(d_o, src_lane) + 8 * __shfl(d_r, src_lane);
        if (pass == 0) {
            if (it < LP_ITEMS && on) {
                const int q = lane - it * LP_LANES, h = q / 5, dx = q - 5 * h - 2;
                const float *o = out4 + (size_t)obase + 2 * h;
                const int hx =
                    clampi((int)rintf(fminf(fmaxf(LP_SCALE_PEL * o[0], -256.0f), 256.0f)), 2 - LP_BOX, LP_BOX - 2);
                const int hy =
                    clampi((int)rintf(fminf(fmaxf(LP_SCALE_PEL * o[1], -256.0f), 256.0f)), 2 - LP_BOX, LP_BOX - 2);
                const int px = hx + dx;
                row0 = hy - 2 + LP_BOX;
                bo = px + LP_BOX + sh;
                idx0 = row0 * LP_BD + px + LP_BOX;
                slot = it;
            } else {
                it = LP_ITEMS;
            }
        } else if (it < LP_ITEMS) {
            const int q = t - (t / LP_FB_TASKS) * LP_FB_TASKS;
            slot = 4 + it;
            if (q < 18) {  // g's +-4 window: column q % 9, rows -4..0 or 1..4
                const int band = q >= 9, px = q - 9 * band - 4;
                row0 = (band ? 1 : -4) + LP_BOX;
                ndy = band ? 4 : 5;
                bo = px + LP_BOX + sh;
                idx0 = row0 * LP_BD + px + LP_BOX;
            } else {  // zero's +-2 window, column wx
                const int wx = q - 20;
                if (abs(gx) <= LP_BOX - 2 && abs(gy) <= LP_BOX - 2) {
                    const int px = wx - gx;
                    row0 = -2 - gy + LP_BOX;
                    bo = px + LP_BOX + sh;
                    idx0 = row0 * LP_BD + px + LP_BOX;
                } else {
                    src_off = zero_off;
                    sp = ly.zp;
                    bo = wx + 2;
                    row0 = 0;
                    zx = wx;
                }
            }
        }
        if (it < LP_ITEMS) {
            const uint32_t *pic = s_w + it * idw;
            const int nd = (xl + 3) >> 2;
            const uint32_t lmask = (xl & 3) ? (1u << (8 * (xl & 3))) - 1u : 0xffffffffu;
            const int key =
                full16 ? lp_column_key<16, 4>(pic, ly.pp, pic + src_off, sp, bo, row0, yl, nd, lmask, ndy, idx0, zx)
                : full24 ? lp_column_key<24, 6>(pic, ly.pp, pic + src_off, sp, bo, row0, yl, nd, lmask, ndy, idx0, zx)
                         : lp_column_key<0, 0>(pic, ly.pp, pic + src_off, sp, bo, row0, yl, nd, lmask, ndy, idx0, zx);
            atomicMin(&s_key[slot], key);
        }
        WAVE_SYNC();
        if (pass == 0) {
            // the winners; items whose SAD exceeds 1.2 x the predicted SAD fall back
            const int key = s_key[me];
            const unsigned long long fl =
                __ballot(lane % LP_LANES == 0 && d_on && (float)(key >> 10) > 1.2f * d_pred);
            fbm = (int)((fl & 1ull) | (((fl >> 20) & 1ull) << 1) | (((fl >> 40) & 1ull) << 2));
            nfb = __popc(fbm);
#pragma unroll
            for (int i = 0; i < LP_ITEMS; ++i) {
                const int ki = U(key, i);
                best_sad[i] = ki >> 10;
                best_idx[i] = ki & 1023;
            }
            if (!nfb) break;
            // zero's window outside the box: staged apart (realigned, clamped if needed)
#pragma unroll 1
            for (int i = 0; i < LP_ITEMS; ++i) {
                if (!((fbm >> i) & 1) || (abs(U(d_gx, i)) <= LP_BOX - 2 && abs(U(d_gy, i)) <= LP_BOX - 2)) continue;
                const JobDesc &J = jobs[U(d_job, i)];
                const int bp = J.p8, r = U(d_r, i), xl = U(d_xl, i), yl = U(d_yl, i);
                const uint8_t *const base[3] = {J.ref8[r], J.ref8[r], J.ref8[r]};
                const int x0[3] = {U(d_xp, i) - 2, 0, 0}, y0[3] = {U(d_yp, i) - 2, 0, 0};
                B8Rect rc[3];
                bool fast[3];
                rc[0] = b8_rect(base[0], bp, s_w + i * idw + zero_off, ly.zp, x0[0], y0[0], xl + 4, yl + 4, xl + 4);
                rc[1] = rc[2] = rc[0];
                fast[0] = b8_inside(bp, J.H, x0[0], y0[0], yl + 4, xl + 4);
                fast[1] = fast[2] = false;
                b8_stage(rc, fast, 1, base, bp, J.W, J.H, x0, y0);
            }
            WAVE_SYNC();
            pass = 1;
            t0 = 0;
        } else {
            t0 += WAVE;
            if (t0 >= nfb * LP_FB_TASKS) break;
        }
    }

    // results, one item at a time; the counters are added once per run of
    // items of one slot
    const int key2 = s_key[4 + me];
    unsigned long long *const cst = stats + (size_t)((xcd_block() * 4 + wv) % kLpStatCopies) * kLpStatStride;
    int cslot = -1, cb = 0, cf = 0;
#pragma unroll 1
    for (int i = 0; i < LP_ITEMS; ++i) {
        if (gbase + i >= total) break;
        const JobDesc &J = jobs[U(d_job, i)];
        const int r = U(d_r, i), k = U(d_k, i);
        if (!((onbits >> i) & 1)) {  // out of the picture: no update (pixel_match.cpp:278-281)
            if (lane == 0) {
                J.vec[r][k].x = J.vec[r][k].y = 0;
                J.pc[r][k].SAD = 10000000.0f;
            }
            continue;
        }
        int sad = i == 0 ? best_sad[0] : (i == 1 ? best_sad[1] : best_sad[2]);
        int idx = i == 0 ? best_idx[0] : (i == 1 ? best_idx[1] : best_idx[2]);
        const bool fbi = (fbm >> i) & 1;
        bool zero_win = false;
        if (fbi) {
            const int k2 = U(key2, i);
            if ((k2 >> 10) < sad) {
                sad = k2 >> 10;
                idx = k2 & 1023;
                zero_win = idx >= LP_BD * LP_BD;
            }
        }
        dme_mv best;
        if (zero_win) {
            best = win_off(idx - LP_BD * LP_BD, 2);
        } else {
            best.x = U(d_gx, i) + idx % LP_BD - LP_BOX;
            best.y = U(d_gy, i) + idx / LP_BD - LP_BOX;
        }
        if (lane == 0) {
            J.vec[r][k] = best;
            J.pc[r][k].SAD = (float)sad;
        }
        const int slot = J.lp_slot;
        if (slot != cslot) {
            if (cslot >= 0 && lane == 0) {
                atomicAdd(cst + (size_t)cslot * kLpStatSlot, (unsigned long long)cb);
                if (cf) atomicAdd(cst + (size_t)cslot * kLpStatSlot + 1, (unsigned long long)cf);
            }
            cslot = slot;
            cb = cf = 0;
        }
        ++cb;
        cf += fbi;
    }
    if (cslot >= 0 && lane == 0) {
        atomicAdd(cst + (size_t)cslot * kLpStatSlot, (unsigned long long)cb);
        if (cf) atomicAdd(cst + (size_t)cslot * kLpStatSlot + 1, (unsigned long long)cf);
    }
}

// Level 0 of the pel search for luma jobs, from the byte planes: per block
// the SADs of the 5 x 5 candidates within R = 2 of zero and of the guide, the
// first minima of the nine radius-1 windows centred within one pel of each
// centre, and the block's record -- what pel_bulk_block<true> computes, laid
// out for throughput: PB_BLOCKS consecutive blocks per wave, ten lanes per
// block (table t, column c) each sliding down its column's five candidates
// (lp_slide), so that a region row is read and realigned once for five SADs.
#define PB_BLOCKS 6
#define PB_LANES 10

// LDS of one wave of k_pel_bulk8 (dwords): per block the picture block
// (prows rows of pp dwords, rows 16-byte aligned), the zero- and
// guide-centred regions (rrows rows of rp each), the two 5 x 5 tables and
// the 18 window keys.
struct Pb8Lds {
    int pp, prows, rp, rrows;
    int ndm, ncm;  // dwords of the widest block row and of the widest staged region row
    __device__ int item() const { return (pp * prows + 2 * rp * rrows + 50 + 18 + 3) & ~3; }
    __device__ int wave() const { return PB_BLOCKS * item(); }
};

__global__ __launch_bounds__(256) void k_pel_bulk8(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                   int ntasks, int total, Pb8Lds ly) {
    extern __shared__ uint32_t s_pb[];
    const int wv = wave_id(), lane = threadIdx.x & (WAVE - 1);
    const int gbase = (xcd_block() * 4 + wv) * PB_BLOCKS;
    if (gbase >= total) return;
    const int idw = ly.item(), reg_off = ly.pp * ly.prows, tab_off = reg_off + 2 * ly.rp * ly.rrows;
    uint32_t *s_w = s_pb + wv * ly.wave();
    // the lanes of group i = min(lane / 10, 5) hold the descriptor of block gbase + i
    const int me = min(lane / PB_LANES, PB_BLOCKS - 1), gi = min(gbase + me, total - 1);
    int cnt = 0;
    for (int j0 = 0; j0 < ntasks; j0 += WAVE) {
        const int j = j0 + lane;
        const int off = j < ntasks ? tasks[j].offset : INT_MAX;
#pragma unroll
        for (int i = 0; i < PB_BLOCKS; ++i) {
            const int c = __popcll(__ballot(off <= min(gbase + i, total - 1)));
            if (i == me) cnt += c;
        }
    }
    const Task tk = tasks[cnt - 1];
    const JobDesc &J = jobs[tk.job];
    const int local = gi - tk.offset, X = J.lxb[0], xb = local % X, yb = local / X;
    const Blk b = block_limits(J.blen[2], J.bsep[2], J.boff[2], J.lw[0], J.lh[0], xb, yb);
    const int has_guide = 0 < J.depth;
    dme_mv g;
    g.x = g.y = 0;
    if (has_guide) g = pel_guide(J, tk.ref, 0, local);
    const int valid = gbase + me < total, on = valid && !b.empty();
    auto U = [&](int v, int i) { return __builtin_amdgcn_readlane(v, i * PB_LANES); };
    const unsigned long long onl = __ballot(lane % PB_LANES == 0 && lane < PB_BLOCKS * PB_LANES && on);
    // the slide's unrolled forms: every active block whole at 16 x 16, 12 x 12 or 24 x 24
    const unsigned long long odd16 = __ballot(on && (b.xl != 16 || b.yl != 16));
    const unsigned long long odd12 = __ballot(on && (b.xl != 12 || b.yl != 12));
    const unsigned long long odd24 = __ballot(on && (b.xl != 24 || b.yl != 24));

    // staging of all blocks at once.  Task (block i, k): k < 2 ND a half
    // (even or odd rows) of a dword column of the picture block, realigned
    // to the block's first column and zero past its width; the other tasks a
    // dword column of the zero- or guide-centred region, copied whole from the
    // dword at or before the region's first column (its shift goes into the
    // slide's byte offset).  Columns near the planes' edges are read sample by
    // sample with clamped coordinates.
    const int ndm = ly.ndm, ncm = ly.ncm, ntask = 2 * ndm + 2 * ncm;
    auto region_x0 = [&](int xp, int gx, int t) { return xp - 2 + (t ? gx : 0); };
    for (int t0 = 0; t0 < PB_BLOCKS * ntask; t0 += WAVE) {
        const int task = t0 + lane, i = min(task / ntask, PB_BLOCKS - 1), k = task - i * ntask;
        const int src = i * PB_LANES;
        const int ion = (int)((onl >> src) & 1), igd = __shfl(has_guide, src);
        const int ijob = __shfl(tk.job, src), ir = __shfl(tk.ref, src);
        const int xp = __shfl(b.xp, src), yp = __shfl(b.yp, src), xl = __shfl(b.xl, src), yl = __shfl(b.yl, src);
        const int gx = __shfl(g.x, src), gy = __shfl(g.y, src);
        if (task >= PB_BLOCKS * ntask || !ion) continue;
        const JobDesc &Ji = jobs[ijob];
        const int bp = Ji.p8, W = Ji.lw[0], H = Ji.lh[0], nd = (xl + 3) >> 2;
        uint32_t *s_pic = s_w + i * idw;
        if (k < 2 * ndm) {
            const int col = k >> 1, half = k & 1;
            if (col >= nd) continue;
            const int valid_b = xl - 4 * col;
            const uint32_t mask = valid_b >= 4 ? 0xffffffffu : ((1u << (8 * valid_b)) - 1u);
            const int sh = xp & 3;
            const bool inside = (xp & ~3) + 4 * col + 8 <= bp - B8_MG && yp + yl <= H + B8_MG;
            const uint32_t *q = reinterpret_cast<const uint32_t *>(Ji.pic8 + (ptrdiff_t)yp * bp + (xp & ~3)) + col;
            for (int r0 = half; r0 < yl; r0 += 16) {
                uint32_t v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int row = min(r0 + 2 * u, yl - 1);  // rows past the block repeat its last row
                    if (inside) {
                        const uint32_t *w = q + (ptrdiff_t)row * (bp >> 2);
                        v[u] = __builtin_amdgcn_alignbyte(w[1], w[0], sh);
                    } else {
                        const uint8_t *sr = Ji.pic8 + (ptrdiff_t)clampi(yp + row, 0, H - 1) * bp;
                        uint32_t x = 0;
#pragma unroll
                        for (int e = 0; e < 4; ++e) x |= (uint32_t)sr[clampi(xp + 4 * col + e, 0, W - 1)] << (8 * e);
                        v[u] = x;
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) s_pic[min(r0 + 2 * u, yl - 1) * ly.pp + col] = v[u] & mask;
            }
        } else {
            const int kk = k - 2 * ndm, t = kk >= ncm, col = kk - t * ncm;
            if (t && !igd) continue;
            const int x0 = region_x0(xp, gx, t), y0 = yp - 2 + (t ? gy : 0), RW = xl + 4, RH = yl + 4;
            const int sh = x0 & 3, ncol = (RW + sh + 3) >> 2;
            if (col >= ncol) continue;
            uint32_t *dst = s_pic + reg_off + t * ly.rp * ly.rrows + col;
            const int ax = (x0 & ~3) + 4 * col;
            const bool inside = ax >= -B8_MG && ax + 4 <= bp - B8_MG && y0 >= -B8_MG && y0 + RH <= H + B8_MG;
            const uint32_t *q = reinterpret_cast<const uint32_t *>(Ji.ref8[ir] + (ptrdiff_t)y0 * bp + ax);
            for (int r0 = 0; r0 < RH; r0 += 12) {
                uint32_t v[12];
#pragma unroll
                for (int u = 0; u < 12; ++u) {
                    const int row = min(r0 + u, RH - 1);
                    if (inside) {
                        v[u] = q[(ptrdiff_t)row * (bp >> 2)];
                    } else {
                        const uint8_t *sr = Ji.ref8[ir] + (ptrdiff_t)clampi(y0 + row, 0, H - 1) * bp;
                        uint32_t x = 0;
#pragma unroll
                        for (int e = 0; e < 4; ++e) x |= (uint32_t)sr[clampi(ax + e, 0, W - 1)] << (8 * e);
                        v[u] = x;
                    }
                }
#pragma unroll
                for (int u = 0; u < 12; ++u) dst[min(r0 + u, RH - 1) * ly.rp] = v[u];
            }
        }
    }
    WAVE_SYNC();
    // the tables: lane (block, t, c) slides down column c of table t
    {
        const int q = lane - me * PB_LANES, t = q / 5, c = q - 5 * t;
        if (lane < PB_BLOCKS * PB_LANES && on && (t == 0 || has_guide)) {
            const uint32_t *pic = s_w + me * idw, *src = pic + reg_off + t * ly.rp * ly.rrows;
            const int nd = (b.xl + 3) >> 2;
            const uint32_t lmask = (b.xl & 3) ? (1u << (8 * (b.xl & 3))) - 1u : 0xffffffffu;
            uint32_t acc[5] = {0u, 0u, 0u, 0u, 0u};
            const int bo = c + ((b.xp - 2 + (t ? g.x : 0)) & 3);
            if (!odd16) lp_slide<16, 4>(pic, ly.pp, src, ly.rp, bo, 0, b.yl, nd, lmask, acc);
            else if (!odd12) lp_slide<12, 3>(pic, ly.pp, src, ly.rp, bo, 0, b.yl, nd, lmask, acc);
            else if (!odd24) lp_slide<24, 6>(pic, ly.pp, src, ly.rp, bo, 0, b.yl, nd, lmask, acc);
            else lp_slide<0, 0>(pic, ly.pp, src, ly.rp, bo, 0, b.yl, nd, lmask, acc);
            int *tab = reinterpret_cast<int *>(s_w + me * idw + tab_off) + t * 25;
#pragma unroll
            for (int dy = 0; dy < 5; ++dy) tab[dy * 5 + c] = (int)acc[dy];
        }
    }
    WAVE_SYNC();
    // the window keys (SAD * 128 + enumeration index) of every block: lane
    // task (block i, t * 9 + w) scans window w of table t
    for (int t0 = 0; t0 < PB_BLOCKS * 18; t0 += WAVE) {
        const int task = t0 + lane, i = min(task / 18, PB_BLOCKS - 1), q = task - i * 18;
        const int ion = (int)((onl >> (i * PB_LANES)) & 1), igd = __shfl(has_guide, i * PB_LANES);
        if (task < PB_BLOCKS * 18) {
            const int t = q / 9, w = q - 9 * t;
            int *tabs = reinterpret_cast<int *>(s_w + i * idw + tab_off);
            int key = INT_MAX;
            if (ion && (t == 0 || igd)) key = pel_window_min<1>(tabs + t * 25 + (w / 3 + 1) * 5 + (w % 3 + 1));
            tabs[50 + q] = key;
        }
    }
    WAVE_SYNC();
    // the records: keys, then the winner of the zero list and of the guide
    // list (radius 1, centred windows)
    if (lane < PB_BLOCKS * PB_LANES && valid) {
        const int q = lane - me * PB_LANES;
        const int *keys = reinterpret_cast<const int *>(s_w + me * idw + tab_off) + 50;
        int *rec = J.ptab[tk.ref] + (size_t)wf_pos(xb, yb, X, J.lyb[0]) * PEL_REC;
        rec[8 + q] = keys[q];
        if (q < 8) rec[18 + q] = keys[10 + q];
        if (q == 0) {
            int best_sad = 10000000;
            dme_mv best_mv;
            best_mv.x = best_mv.y = 0;
            if (on) {
                const int k0 = keys[4], k1 = keys[13];
                best_sad = k0 >> 7;
                best_mv = win_off(k0 & 127, 1);
                if (has_guide && (k1 >> 7) < best_sad) {
                    best_sad = k1 >> 7;
                    best_mv = win_off(k1 & 127, 1);
                    best_mv.x += g.x;
                    best_mv.y += g.y;
                }
            }
            rec[0] = best_sad;
            rec[1] = best_mv.x;
            rec[2] = best_mv.y;
            rec[3] = g.x;
            rec[4] = g.y;
        }
    }
}

// Pel costs of the proposed vectors: the prediction rules of the pel search
// (pixel_match.cpp:314-347) over the final level-0 vectors, mvcost =
// Norm1(mv - pred) through its short, total = SAD (SetTotal(0.0)).
__global__ __launch_bounds__(256) void k_lp_pel_costs(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                      int ntasks, int total) {
    const int gi = blockIdx.x * 256 + threadIdx.x;
    if (gi >= total) return;
    const Task tk = tasks[find_task(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    const int k = gi - tk.offset, X = J.xnb, xb = k % X, yb = k / X;
    const dme_mv *__restrict__ V = J.vec[tk.ref];
    dme_mv pred;
    pred.x = pred.y = 0;
    if (yb == 0) {
        if (xb > 0) pred = V[k - 1];
    } else if (xb == 0) {
        pred = V[k - X];
    } else if (xb < X - 1) {
        pred = med3(V[k - 1], V[k - X], V[k - X + 1]);
    } else {
        pred = mean2(V[k - X], V[k - 1]);
    }
    const dme_mv v = V[k];
    dme_cost c = J.pc[tk.ref][k];
    c.mvcost = (float)(int16_t)norm1(v.x - pred.x, v.y - pred.y);
    c.total = c.SAD + 0.0f * c.mvcost;  // SetTotal(0.0)
    J.pc[tk.ref][k] = c;
}

// ---------------------------------------------------------------------------
// K7: ModeDecider::SetDC (me_mode_decn.cpp:523-573): block DCs of Y, U, V,
// then the 3x3 [1 3 1; 3 0 3; 1 3 1]/16 smoothing; row 0 and the first and
// last columns are copied and the last row is left as it was.
// ---------------------------------------------------------------------------
// SetDC's block means (me_mode_decn.cpp:540-550, GetBlockDC :523-538): one
// workgroup per (picture, component, block row).  Every block of a row covers
// the same picture rows, so the workgroup first sums each column over those
// rows (coalesced across threads) into LDS, then each block adds its columns'
// sums.  Integer sums are order-independent: the same value as the
// reference's row-by-row loop.  Blocks clipped away entirely get 0, as do
// chroma blocks of a picture registered without chroma.
__global__ __launch_bounds__(256) void k_block_dc(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                  int ntasks, int total) {
    extern __shared__ int s_col[];  // one int per column of the widest plane
    const int row = blockIdx.x;
    if (row >= total) return;
    const Task tk = tasks[find_task_wave(tasks, ntasks, row)];
    const JobDesc &J = jobs[tk.job];
    const int comp = tk.ref, yb = row - tk.offset;
    const int16_t *__restrict__ p = comp == 0 ? J.py : (comp == 1 ? J.cu : J.cv);  // SetDC reads luma
    const int *blen = comp == 0 ? J.blen[2] : J.cblen;
    const int *bsep = comp == 0 ? J.bsep[2] : J.cbsep;
    const int *boff = comp == 0 ? J.boff[2] : J.cboff;
    const int W = comp == 0 ? J.W : J.CW, H = comp == 0 ? J.H : J.CH;
    const Blk rb = block_limits(blen, bsep, boff, W, H, 0, yb);  // the row range of every block in the row
    const bool any = p && rb.yl > 0;
    if (any) {
        for (int c = threadIdx.x; c < W; c += 256) {
            const int16_t *q = p + (size_t)rb.yp * W + c;
            int sum = 0;
            for (int r = 0; r < rb.yl; ++r) sum += q[(size_t)r * W];
            s_col[c] = sum;
        }
    }
    __syncthreads();
    for (int xb = threadIdx.x; xb < J.xnb; xb += 256) {
        const Blk b = block_limits(blen, bsep, boff, W, H, xb, yb);
        int dc = 0;
        if (any && !b.empty()) {
            int sum = 0;
            for (int c = 0; c < b.xl; ++c) sum += s_col[b.xp + c];
            dc = (int16_t)(sum / (b.xl * b.yl));
        }
        J.dct[comp][yb * J.xnb + xb] = (int16_t)dc;
    }
}

__global__ __launch_bounds__(256) void k_dc_smooth(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                   int ntasks, int total) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= total) return;
    const Task tk = tasks[find_task(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    const int comp = tk.ref, k = gi - tk.offset;
    const int X = J.xnb, Y = J.ynb;
    const int x = k % X, y = k / X;
    if (y == Y - 1) return;
    if (comp > 0 && !(comp == 1 ? J.cu : J.cv)) return;
    const int16_t *t = J.dct[comp];
    int v;
    if (y == 0 || x == 0 || x == X - 1) {
        v = t[k];
    } else {
        v = (t[k - X - 1] + 3 * t[k - X] + t[k - X + 1] + 3 * t[k - 1] + 3 * t[k + 1] + t[k + X - 1] + 3 * t[k + X] +
             t[k + X + 1] + 8) >> 4;
    }
    J.dc[comp][k] = (int16_t)v;
}

// ---------------------------------------------------------------------------
// Launch wrappers (host side, same translation unit as the kernels)
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// The fast form of the learned path (no counterpart in the reference).  The
// learned stage gives each block its level-0 integer vector; then every
// decision is made without waiting for its neighbours' decisions of the same
// pass, and the reference's Lagrangian of the final vectors is evaluated with
// the reference's own predictors:
//   k_fast_subpel  RefineMatchSubp (block_match.cpp:327-467) with the
//                  predictor of the integer vectors (GetPred of me_subpel.cpp
//                  on them), costs float(SAD) + start in one rounding
//   k_fast_l2      the level-2 costs SubpelRefine leaves, with GetPred of the
//                  final sub-pel vectors
//   k_mode_l2      (shared with the plain path) level-2 unit decisions
//   k_fast_mode    levels 1 and 0 of every superblock with DoME's predictor
//                  taken from the pre-merge level-2 vectors; the split
//   k_fast_eval    superblocks split into units: the unit costs with the
//                  predictor DoModeDecn's raster pass would form (the final
//                  vectors of earlier superblocks, the pre-merge ones inside),
//                  DoUnitDecn and the sum
// The oracle restates the same steps (oracle_fast_me).
// ---------------------------------------------------------------------------

// Packed biased bytes of samples x .. x + 3 of row y of a byte plane (origin
// dword-aligned, margins around the picture).
__device__ __forceinline__ uint32_t b8_dword(const uint8_t *__restrict__ origin, int pitch, int x, int y) {
    const uintptr_t a = (uintptr_t)(origin + (ptrdiff_t)y * pitch + x);
    const uint32_t *q = reinterpret_cast<const uint32_t *>(a & ~(uintptr_t)3);
    return __builtin_amdgcn_alignbyte(q[1], q[0], (uint32_t)(a & 3));
}

// Byte up-planes (the fast form's references, k_build_qp<true>).  A plane
// holds the reference's up-converted samples (UpConverter, upconvert.cpp:
// 55-200) as biased bytes, its margin replicating the edge samples, so that
// a read inside the plane is the reference's clamped read (BChk,
// common.h:1574-1579).  A quarter-pel sample (BlockDiffQuarterPel /
// BiBlockQuarterPel, me_utils.cpp:573-867, 1288-1543) at up position
// (ux + rx / 2, uy + ry / 2) is a, (a + b + 1) >> 1, (a + c + 1) >> 1 or
// (a + b + c + d + 2) >> 2 of the up samples a (ux, uy), b (ux + 1, uy),
// c (ux, uy + 1) and d (ux + 1, uy + 1); the bias of 128 commutes with all
// four.

// Clamped read of up sample (ux, uy) of reference r (biased).
__device__ __forceinline__ int up8_at(const JobDesc &J, int r, int ux, int uy) {
    return J.up8[r][(ptrdiff_t)clampi(uy, 0, 2 * J.H - 1) * J.u8p + clampi(ux, 0, 2 * J.W - 1)];
}

// Quarter-pel sample (biased) at quarter-pel position (qx, qy), clamped reads.
__device__ __forceinline__ int up8_sample(const JobDesc &J, int r, int qx, int qy) {
    const int ux = qx >> 1, uy = qy >> 1, rx = qx & 1, ry = qy & 1;
    const int a = up8_at(J, r, ux, uy), b = up8_at(J, r, ux + 1, uy);
    const int c = up8_at(J, r, ux, uy + 1), d = up8_at(J, r, ux + 1, uy + 1);
    return ((2 - rx) * (2 - ry) * a + rx * (2 - ry) * b + (2 - rx) * ry * c + rx * ry * d + 2) >> 2;
}

// Rounded averages of packed bytes: (a + b + 1) >> 1 (v_lerp_u8) and
// (a + b + c + d + 2) >> 2 (in 16-bit lanes).
__device__ __forceinline__ uint32_t avg2_u8(uint32_t a, uint32_t b) { return __builtin_amdgcn_lerp(a, b, 0x01010101u); }
__device__ __forceinline__ uint32_t avg4_u8(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    const uint32_t m = 0x00ff00ffu;
    const uint32_t lo = (a & m) + (b & m) + (c & m) + (d & m) + 0x00020002u;
    const uint32_t hi = ((a >> 8) & m) + ((b >> 8) & m) + ((c >> 8) & m) + ((d >> 8) & m) + 0x00020002u;
    return ((lo >> 2) & m) | (((hi >> 2) & m) << 8);
}

// The even and the odd bytes of the 8 bytes from byte offset o of the dwords
// at q (reads the three dwords from dword o / 4).  Offsets, not integer
// casts of pointers, so that LDS reads stay LDS reads.
__device__ __forceinline__ void bytes8(const uint32_t *__restrict__ q, uint32_t o, uint32_t &even, uint32_t &odd) {
    const uint32_t *__restrict__ p = q + (o >> 2);
    const uint32_t sh = o & 3u;
    const uint32_t d0 = __builtin_amdgcn_alignbyte(p[1], p[0], sh), d1 = __builtin_amdgcn_alignbyte(p[2], p[1], sh);
    even = __builtin_amdgcn_perm(d1, d0, 0x06040200u);
    odd = __builtin_amdgcn_perm(d1, d0, 0x07050301u);
}

// Packed quarter-pel samples of four consecutive pixels of one phase (rx,
// ry) whose up samples start at byte offset o of q (the next up row `pitch`
// bytes further).
__device__ __forceinline__ uint32_t quad_u8(const uint32_t *__restrict__ q, uint32_t o, uint32_t pitch, int rx,
                                            int ry) {
    uint32_t a, b;
    bytes8(q, o, a, b);
    if (!ry) return rx ? avg2_u8(a, b) : a;
    uint32_t c, d;
    bytes8(q, o + pitch, c, d);
    return rx ? avg4_u8(a, b, c, d) : avg2_u8(a, c);
}

// Lane layout of a wave over a unit of xl x yl samples: dword column col
// (four samples) of rows row0, row0 + rpp, ...; mask keeps the samples of a
// partial last column.
struct UnitLanes {
    int nd, rpp, row0, col, npass;
    uint32_t mask;
    bool on;
};

__device__ __forceinline__ UnitLanes unit_lanes(const Blk &b) {
    UnitLanes L;
    const int lane = threadIdx.x & (WAVE - 1);
    L.nd = (b.xl + 3) >> 2;
    L.rpp = WAVE / L.nd;
    L.row0 = (int)(((float)lane + 0.5f) * (1.0f / (float)L.nd));
    L.col = lane - L.row0 * L.nd;
    L.on = L.row0 < L.rpp;
    const int valid = b.xl - 4 * L.col;
    L.mask = valid >= 4 ? 0xffffffffu : ((1u << (8 * valid)) - 1u);
    L.npass = (b.yl + L.rpp - 1) / L.rpp;
    return L;
}

// Row pass p of lane layout L: the row, or -1 off the unit.
__device__ __forceinline__ int unit_row(const UnitLanes &L, int p, int yl) {
    const int row = L.row0 + p * L.rpp;
    return L.on && row < yl ? row : -1;
}

// A quarter-pel candidate m of unit b against reference r's byte up-plane:
// lane (col, row)'s four pixels start at up column 2 (xp + 4 col) + (m.x >> 1)
// of up row 2 (yp + row) + (m.y >> 1).  `fast` when every read of the unit
// lies inside the plane (its margin replicates the clamped samples);
// otherwise the samples are read one by one with clamped coordinates.
struct Up8Cand {
    const uint32_t *plane;  // the plane's first dword (up row and column -kUp8Margin)
    uint32_t o0;            // byte offset of the unit's first up sample
    int rx, ry, qx0, qy0, r;
    bool fast;
};

__device__ __forceinline__ Up8Cand up8_cand(const JobDesc &J, int r, const Blk &b, int nd, dme_mv m) {
    Up8Cand c;
    c.r = r;
    c.rx = m.x & 1;
    c.ry = m.y & 1;
    const int ux0 = 2 * b.xp + (m.x >> 1), uy0 = 2 * b.yp + (m.y >> 1);
    c.plane = reinterpret_cast<const uint32_t *>(J.up8[r] - ((ptrdiff_t)kUp8Margin * J.u8p + kUp8Margin));
    c.fast = ux0 >= -kUp8Margin && ux0 + 8 * nd + 12 <= J.u8p - kUp8Margin && uy0 >= -kUp8Margin &&
             uy0 + 2 * b.yl + 2 <= up8_rows(J.H) - kUp8Margin;
    c.o0 = c.fast ? (uint32_t)((uy0 + kUp8Margin) * J.u8p + ux0 + kUp8Margin) : 0u;
    c.qx0 = 4 * b.xp + m.x;
    c.qy0 = 4 * b.yp + m.y;
    return c;
}

__device__ __forceinline__ uint32_t up8_pack(const JobDesc &J, const Up8Cand &c, int col, int row) {
    if (c.fast) return quad_u8(c.plane, c.o0 + (uint32_t)(2 * row * J.u8p + 8 * col), (uint32_t)J.u8p, c.rx, c.ry);
    uint32_t v = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        v |= (uint32_t)up8_sample(J, c.r, c.qx0 + 4 * (4 * col + i), c.qy0 + 4 * row) << (8 * i);
    return v;
}

// Up to four wave-uniform candidates and their SADs, in registers.
struct Cand4 {
    dme_mv c[4];
};
struct Sad4 {
    int s[4];
};

// Integer SADs of the first n (<= N) quarter-pel candidates of unit b
// against reference r; the unit's packed picture dwords are
// s_pic[row * pd + col].  Wave-uniform results.
template <int N>
__device__ __forceinline__ Sad4 up8_sads(const JobDesc &J, int r, const Blk &b, const UnitLanes &L,
                                         const uint32_t *__restrict__ s_pic, int pd, const Cand4 &c, int n) {
    Up8Cand q[N];
#pragma unroll
    for (int i = 0; i < N; ++i) q[i] = up8_cand(J, r, b, L.nd, c.c[i < n ? i : 0]);
    uint32_t acc[N];
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = 0u;
    for (int p = 0; p < L.npass; ++p) {
        const int row = unit_row(L, p, b.yl);
        if (row < 0) continue;
        const uint32_t pv = s_pic[row * pd + L.col];
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (i < n) acc[i] = __builtin_amdgcn_sad_u8(pv, up8_pack(J, q[i], L.col, row) & L.mask, acc[i]);
    }
    Sad4 out;
#pragma unroll
    for (int i = 0; i < 4; ++i) out.s[i] = (i < N && i < n) ? wave_sum((int)acc[i < N ? i : 0]) : 0;
    return out;
}

// Bi-prediction SAD of unit b (BiBlock*Pel::Diff, me_utils.cpp:1172-1543):
// |((2p - r1) - r2) >> 1| = |p - ceil((r1 + r2) / 2)|, a byte SAD against the
// rounded-up average of the two candidates' biased samples.  Vectors in
// 1/2^step pel.
__device__ __forceinline__ int up8_bi_sad_s(const JobDesc &J, const Blk &b, const UnitLanes &L,
                                            const uint32_t *__restrict__ s_pic, int pd, dme_mv c1, dme_mv c2) {
    const int qs = 2 - J.step;
    c1.x <<= qs;
    c1.y <<= qs;
    c2.x <<= qs;
    c2.y <<= qs;
    const Up8Cand q1 = up8_cand(J, 0, b, L.nd, c1), q2 = up8_cand(J, 1, b, L.nd, c2);
    uint32_t acc = 0;
    for (int p = 0; p < L.npass; ++p) {
        const int row = unit_row(L, p, b.yl);
        if (row < 0) continue;
        const uint32_t a = up8_pack(J, q1, L.col, row), d = up8_pack(J, q2, L.col, row);
        acc = __builtin_amdgcn_sad_u8(s_pic[row * pd + L.col], avg_up_u8(a, d) & L.mask, acc);
    }
    return wave_sum((int)acc);
}

// The same for a level-2 block without a staged picture (k_mode_l2<true>).
__device__ int up8_bi_sad(const JobDesc &J, const Blk &b, dme_mv m1, dme_mv m2) {
    const UnitLanes L = unit_lanes(b);
    const int qs = 2 - J.step;
    m1.x <<= qs;
    m1.y <<= qs;
    m2.x <<= qs;
    m2.y <<= qs;
    const Up8Cand q1 = up8_cand(J, 0, b, L.nd, m1), q2 = up8_cand(J, 1, b, L.nd, m2);
    uint32_t acc = 0;
    for (int p = 0; p < L.npass; ++p) {
        const int row = unit_row(L, p, b.yl);
        if (row < 0) continue;
        const uint32_t pv = b8_dword(J.pic8, J.p8, b.xp + 4 * L.col, b.yp + row) & L.mask;
        const uint32_t a = up8_pack(J, q1, L.col, row), d = up8_pack(J, q2, L.col, row);
        acc = __builtin_amdgcn_sad_u8(pv, avg_up_u8(a, d) & L.mask, acc);
    }
    return wave_sum((int)acc);
}

// The reference's per-sample float chain of a bounds-checked up-sampled Diff
// (as wave_chain) from reference r's byte up-plane.
__device__ __noinline__ float up8_chain(float start, const JobDesc *__restrict__ jp, int r, Blk b, UpStart st) {
    const JobDesc &J = *jp;
    const int lane = threadIdx.x & (WAVE - 1);
    const int n = b.xl * b.yl;
    const int16_t *__restrict__ pic = J.pic[0];
    float s = start;
    for (int b0 = 0; b0 < n; b0 += 4 * WAVE) {
        int d[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = b0 + k * WAVE + lane;
            const int ii = i < n ? i : n - 1;
            const int dy = ii / b.xl, dx = ii - dy * b.xl;
            const int q = up8_sample(J, r, 2 * st.rsx + st.rx + 4 * dx, 2 * st.rsy + st.ry + 4 * dy) - 128;
            const int v = abs(q - (int)pic[(size_t)(b.yp + dy) * J.W + b.xp + dx]);
            d[k] = i < n ? v : 0;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (b0 + k * WAVE >= n) break;
            int x = d[k];
#pragma unroll
            for (int off = 1; off < WAVE; off <<= 1) {
                const int y = __shfl_up(x, off, WAVE);
                if (lane >= off) x += y;
            }
            const int total = __builtin_amdgcn_readlane(x, WAVE - 1);
            int consumed = 0, first = 0;
            while (first < WAVE) {
                if (s >= 16777216.0f) {
                    for (int j = first; j < WAVE; ++j)
                        s += (float)(__builtin_amdgcn_readlane(x, j) - (j ? __builtin_amdgcn_readlane(x, j - 1) : 0));
                    break;
                }
                const double D = (double)binade_limit(s) - (double)s;
                const unsigned long long m = __ballot(lane >= first && (double)(x - consumed) >= D);
                if (!m) {
                    s += (float)(total - consumed);
                    break;
                }
                const int j = __ffsll((long long)m) - 1;
                const int before = j ? __builtin_amdgcn_readlane(x, j - 1) : 0;
                const int at = __builtin_amdgcn_readlane(x, j);
                s += (float)(before - consumed);
                s += (float)(at - before);
                consumed = at;
                first = j + 1;
            }
        }
    }
    return s;
}

// Stages unit b's packed picture bytes (masked past xl) into s_pic rows of pd
// dwords.
__device__ __forceinline__ void stage_unit_pic(const JobDesc &J, const Blk &b, const UnitLanes &L, uint32_t *s_pic,
                                               int pd) {
    for (int p = 0; p < L.npass; ++p) {
        const int row = unit_row(L, p, b.yl);
        if (row >= 0) s_pic[row * pd + L.col] = b8_dword(J.pic8, J.p8, b.xp + 4 * L.col, b.yp + row) & L.mask;
    }
    WAVE_SYNC();
}

// GetPred (me_subpel.cpp:145-176): MvMedian of the available neighbours
// left, top-left, top, top-right of field V (scaled by `sh`).  Which are
// available depends only on the position: none or the left one in row 0,
// top and top-right in column 0, left, top-left and top in the last column,
// all four elsewhere.
__device__ __forceinline__ dme_mv subpel_pred(const dme_mv *__restrict__ V, int X, int Y, int x, int y, int sh) {
    auto at = [&](int xx, int yy) {
        dme_mv v = V[(size_t)yy * X + xx];
        v.x <<= sh;
        v.y <<= sh;
        return v;
    };
    dme_mv p;
    p.x = p.y = 0;
    if (y == 0) {
        if (x > 0) p = at(x - 1, 0);
    } else if (x == 0) {
        p = X > 1 ? mean2(at(0, y - 1), at(1, y - 1)) : at(0, y - 1);
    } else if (x == X - 1) {
        p = med3(at(x - 1, y), at(x - 1, y - 1), at(x, y - 1));
    } else {
        const dme_mv a = at(x - 1, y), b = at(x - 1, y - 1), c = at(x, y - 1), d = at(x + 1, y - 1);
        p.x = (a.x + b.x + c.x + d.x - max(max(a.x, b.x), max(c.x, d.x)) - min(min(a.x, b.x), min(c.x, d.x)) + 1) >> 1;
        p.y = (a.y + b.y + c.y + d.y - max(max(a.y, b.y), max(c.y, d.y)) - min(min(a.y, b.y), min(c.y, d.y)) + 1) >> 1;
    }
    return p;
}

// Sub-pel refinement of every block at once: RefineMatchSubp's staged
// search (block_match.cpp:327-467) with the predictor of the integer
// vectors.  One wave per (block, reference).  Every candidate of the staged
// search lies within 3 quarter pels of 4 x the integer vector v, so the wave
// first copies the byte up-plane samples those candidates read -- up rows
// 2 (yp + v.y) - 2 .. 2 (yp + v.y + yl) and columns from the dword at or
// before 2 (xp + v.x) - 2 -- into LDS, rows of `pitch` dwords; a candidate's
// four samples of a lane are then one quad_u8 of the staged rows.  The
// predictor, a whole number of pels away, is staged when it equals 4 v and
// read from the plane otherwise.  Writes the integer vectors (fv0), the
// sub-pel vectors (fv2) and the SAD each sub-pel vector's cost is measured
// by (fs2): the pel search's SAD while the vector is the scaled integer
// vector (RefineMatchSubp keeps that cost), else its up-sampled SAD.
struct FsLds {
    int rows;   // staged up rows (2 by + 3)
    int pitch;  // dwords per staged row (2 ceil(bx / 4) + 3, odd)
    int pd;     // dwords per picture row
    int wave;   // dwords per wave: the staged rows, then the picture block
};

// Sum over the wave of two non-negative values below 2^16 in one reduction.
__device__ __forceinline__ void wave_sum2(int &a, int &b, bool packed) {
    if (packed) {
        const int s = wave_sum(a | (b << 16));
        a = s & 0xffff;
        b = (int)((unsigned)s >> 16);
    } else {
        a = wave_sum(a);
        b = wave_sum(b);
    }
}

__global__ __launch_bounds__(256) void k_fast_subpel(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                     int ntasks, int total, FsLds ly) {
    extern __shared__ uint32_t s_fs[];
    const int wv = wave_id(), lane = threadIdx.x & (WAVE - 1);
    const int gi = xcd_block() * 4 + wv;
    if (gi >= total) return;
    uint32_t *s_up = s_fs + wv * ly.wave;
    uint32_t *s_pic = s_up + ly.rows * ly.pitch;
    const Task tk = tasks[find_task_wave(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    const int r = tk.ref, k = gi - tk.offset, X = J.xnb, Y = J.ynb, xb = k % X, yb = k / X;
    const int prec = J.prec, W = J.W, H = J.H;
    const dme_mv *__restrict__ V0 = J.vec[r];
    const dme_mv pel = V0[k];
    const dme_mv pred = subpel_pred(V0, X, Y, xb, yb, prec);
    const Blk b = block_limits(J.blen[2], J.bsep[2], J.boff[2], W, H, xb, yb);
    dme_mv out = pred;
    int out_sad = 0;
    dme_mv pel_s;
    pel_s.x = pel.x << prec;
    pel_s.y = pel.y << prec;
    const float pel_sad = J.pc[r][k].SAD;
    if (!b.empty()) {
        const float lambda = J.lam[k];
        const float lim = (float)(2 * b.xl * b.yl);
        float best_total = pel_sad + lambda * var_up(pred, pel_s);
        if (pel_sad < lim) {
            out = pel_s;
            out_sad = (int)pel_sad;
        } else {
            const UnitLanes L = unit_lanes(b);
            // stage: up rows uy0 .. uy0 + nrow - 1, dwords from up column cx
            const int ux0 = 2 * (b.xp + pel.x) - 2, uy0 = 2 * (b.yp + pel.y) - 2;
            const int cx = ux0 & ~3, a0 = ux0 - cx, nrow = 2 * b.yl + 3, nd = ly.pitch;
            {
                const int rpp = WAVE / nd, r0 = lane / nd, c = lane - r0 * nd;
                const bool inside = cx >= -kUp8Margin && cx + 4 * nd <= J.u8p - kUp8Margin && uy0 >= -kUp8Margin &&
                                    uy0 + nrow <= up8_rows(H) - kUp8Margin;
                if (inside) {
                    // plane rows are whole dwords and the origin is dword-aligned
                    const uint32_t *__restrict__ src =
                        reinterpret_cast<const uint32_t *>(J.up8[r] + (ptrdiff_t)uy0 * J.u8p + cx) + c;
                    const int dp = J.u8p >> 2;
                    // loads of up to eight row passes in flight before any is stored
                    for (int rb = r0 < rpp ? r0 : nrow; rb < nrow; rb += 8 * rpp) {
                        uint32_t d[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const int rr = rb + u * rpp;
                            d[u] = rr < nrow ? src[(ptrdiff_t)rr * dp] : 0u;
                        }
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const int rr = rb + u * rpp;
                            if (rr < nrow) s_up[rr * nd + c] = d[u];
                        }
                    }
                } else {
                    // near the plane's edges: sample by sample, clamped
                    for (int rr = r0 < rpp ? r0 : nrow; rr < nrow; rr += rpp) {
                        uint32_t v = 0;
#pragma unroll
                        for (int i = 0; i < 4; ++i) v |= (uint32_t)up8_at(J, r, cx + 4 * c + i, uy0 + rr) << (8 * i);
                        s_up[rr * nd + c] = v;
                    }
                }
            }
            // per pass (at most four: blocks up to 32 rows, eight or more rows a pass)
            // a lane's picture dword and the byte offset of its samples for offset 0
            uint32_t pv[4];
            int loff[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const int row = p < L.npass ? unit_row(L, p, b.yl) : -1;
                pv[p] = row >= 0 ? b8_dword(J.pic8, J.p8, b.xp + 4 * L.col, b.yp + row) & L.mask : 0u;
                loff[p] = row >= 0 ? (2 * row + 2) * 4 * nd + a0 + 8 * L.col + 2 : -1;
                if (row >= 0) s_pic[row * ly.pd + L.col] = pv[p];  // for the predictor read from the plane
            }
            WAVE_SYNC();
            const bool packed = b.xl * b.yl * 255 < 65536;
            // partial SAD of a lane for quarter-pel offset (dx, dy) from 4 v: up
            // offset (dx >> 1, dy >> 1), phase (dx & 1, dy & 1), the same for every lane
            auto part = [&](int dx, int dy) -> int {
                const int base = (dy >> 1) * 4 * nd + (dx >> 1), rx = dx & 1, ry = dy & 1;
                uint32_t acc = 0;
#pragma unroll
                for (int pp = 0; pp < 4; ++pp) {
                    if (pp >= L.npass) break;
                    if (loff[pp] < 0) continue;
                    // loff + base >= 0: the staged rows start 2 up rows and columns before offset 0
                    const uint32_t v = quad_u8(s_up, (uint32_t)(loff[pp] + base), 4 * nd, rx, ry);
                    acc = __builtin_amdgcn_sad_u8(pv[pp], v & L.mask, acc);
                }
                return (int)acc;
            };
            int ps;
            if (pred.x << (2 - prec) == 4 * pel.x && pred.y << (2 - prec) == 4 * pel.y) {
                int z = 0;
                ps = part(0, 0);
                wave_sum2(ps, z, packed);
            } else {
                Cand4 c;
                c.c[0].x = pred.x << (2 - prec);
                c.c[0].y = pred.y << (2 - prec);
                ps = up8_sads<1>(J, r, b, L, s_pic, ly.pd, c, 1).s[0];
            }
            const float pred_total = (float)ps;
            if (pred_total < lim) {
                out = pred;
                out_sad = ps;
            } else {
                dme_mv best = pel;  // in 1/2^i pel at stage i
                int best_sad = -1;  // still the pel vector
                bool bailed = false;
                for (int i = 1; i <= prec; ++i) {
                    best.x <<= 1;
                    best.y <<= 1;
                    dme_mv tb = best;
                    const int sh = prec - i, q = 2 - i;
                    // offsets from 4 v in quarter pel of candidate c (1/2^i pel)
                    auto qoff = [&](dme_mv c, int &dx, int &dy) {
                        dx = (c.x << q) - 4 * pel.x;
                        dy = (c.y << q) - 4 * pel.y;
                    };
                    auto consider = [&](dme_mv c, int sad) {
                        dme_mv cp;
                        cp.x = c.x << sh;
                        cp.y = c.y << sh;
                        const float t = (float)sad + var_up(pred, cp) * lambda;
                        if (t < best_total) {
                            best_total = t;
                            best_sad = sad;
                            tb = c;
                        }
                    };
                    dme_mv c4[4];
                    c4[0].x = best.x - 1; c4[0].y = best.y;
                    c4[1].x = best.x + 1; c4[1].y = best.y;
                    c4[2].x = best.x;     c4[2].y = best.y - 1;
                    c4[3].x = best.x;     c4[3].y = best.y + 1;
                    int s4[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        int dx, dy;
                        qoff(c4[j], dx, dy);
                        s4[j] = part(dx, dy);
                    }
                    wave_sum2(s4[0], s4[1], packed);
                    wave_sum2(s4[2], s4[3], packed);
#pragma unroll
                    for (int j = 0; j < 4; ++j) consider(c4[j], s4[j]);
                    if (tb.x != best.x || tb.y != best.y) {
                        const bool movx = tb.x != best.x;  // else y moved
                        dme_mv c2[2];
                        c2[0].x = movx ? tb.x : tb.x - 1; c2[0].y = movx ? tb.y - 1 : tb.y;
                        c2[1].x = movx ? tb.x : tb.x + 1; c2[1].y = movx ? tb.y + 1 : tb.y;
                        int s2[2];
#pragma unroll
                        for (int j = 0; j < 2; ++j) {
                            int dx, dy;
                            qoff(c2[j], dx, dy);
                            s2[j] = part(dx, dy);
                        }
                        wave_sum2(s2[0], s2[1], packed);
                        consider(c2[0], s2[0]);
                        consider(c2[1], s2[1]);
                    }
                    best = tb;
                    if ((double)best_total > 1.1 * (double)pred_total) {
                        bailed = true;
                        break;
                    }
                }
                if (bailed) {
                    out = pred;
                    out_sad = ps;
                } else {
                    out = best;
                    out_sad = best_sad < 0 ? (int)pel_sad : best_sad;
                }
            }
        }
        // a vector equal to the scaled integer vector is costed by the pel SAD
        if (out.x == pel_s.x && out.y == pel_s.y) out_sad = (int)pel_sad;
    }
    if (lane == 0) {
        J.fv0[r][k] = pel;
        J.fv2[r][k] = out;
        J.fs2[r][k] = out_sad;
    }
}

// The level-2 costs of the final sub-pel vectors (the PredCosts SubpelRefine
// leaves): SAD, mvcost = GetVarUp(GetPred of the final field, mv), total =
// SetTotal(lambda); {0, 0, 0} for a block with no samples.  One thread per
// (block, reference); the vectors go to the MEData.
__global__ __launch_bounds__(256) void k_fast_l2(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                 int ntasks, int total) {
    const int gi = blockIdx.x * 256 + threadIdx.x;
    if (gi >= total) return;
    const Task tk = tasks[find_task(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    const int r = tk.ref, k = gi - tk.offset, X = J.xnb, Y = J.ynb, xb = k % X, yb = k / X;
    const Blk b = block_limits(J.blen[2], J.bsep[2], J.boff[2], J.W, J.H, xb, yb);
    const dme_mv *__restrict__ V2 = J.fv2[r];
    const dme_mv v = V2[k];
    dme_cost c;
    if (b.empty()) {
        c.SAD = c.mvcost = c.total = 0.0f;
    } else {
        c.SAD = (float)J.fs2[r][k];
        c.mvcost = var_up(subpel_pred(V2, X, Y, xb, yb, 0), v);
        c.total = c.SAD + J.lam[k] * c.mvcost;
    }
    J.vec[r][k] = v;
    J.pc[r][k] = c;
}

// One unit of DoME at level 1 or 0 for one reference
// (me_mode_decn.cpp:279-364, FindBestMatchSubp block_match.cpp:277-325):
// candidates g[0..4) deduplicated in order, the predictor pred, costs
// float(SAD) + start in one rounding, strict < from {0, 0, 1e8} and (0, 0)
// (the zero-initialised level-1/0 MEData).  Returns the best vector and its
// cost and integer SAD.
struct UnitBest {
    dme_mv v;
    dme_cost c;
    int sad;
};

__device__ __forceinline__ UnitBest unit_best(const JobDesc &J, int r, const Blk &b,
                                              const UnitLanes &L, const uint32_t *__restrict__ s_pic, int pd,
                                              const Cand4 &g, dme_mv pred, float lambda) {
    UnitBest u;
    u.v.x = u.v.y = 0;
    u.c.SAD = u.c.mvcost = 0.0f;
    u.c.total = 100000000.0f;
    u.sad = 0;
    if (b.empty()) return u;
    // AddNewVlist's deduplication: the distinct candidates in order (unrolled
    // selects, so that they stay in registers)
    Cand4 cc = g;
    int n = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        bool dup = false;
#pragma unroll
        for (int j = 0; j < i; ++j) dup = dup || (g.c[j].x == g.c[i].x && g.c[j].y == g.c[i].y);
        if (!dup) {
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (t == n) cc.c[t] = g.c[i];
            ++n;
        }
    }
    Cand4 qc;
    const int qs = 2 - J.step;  // mode-decision vectors are in 1/2^step pel
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        qc.c[i].x = cc.c[i].x << qs;
        qc.c[i].y = cc.c[i].y << qs;
    }
    const Sad4 sads = up8_sads<4>(J, r, b, L, s_pic, pd, qc, n);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= n) break;
        const float mvcost = var_up(pred, cc.c[i]);
        const float start = mvcost * lambda;
        const float t = (float)sads.s[i] + start;
        if (t < u.c.total) {
            u.v = cc.c[i];
            u.c.total = t;
            u.c.mvcost = mvcost;
            u.c.SAD = t - start;
            u.sad = sads.s[i];
        }
    }
    return u;
}

// DoUnitDecn below level 2 (me_mode_decn.cpp:368-462): REF1, REF2 by strict
// <, then the bi-prediction; no intra.  Returns the unit cost and its mode.
__device__ __forceinline__ float unit_decn(int nrefs, float lf, float lambda, const dme_cost &c1, const dme_cost &c2,
                                           int bi_sad, int &mode) {
    mode = DME_REF1_ONLY;
    float best = c1.total * lf;
    if (nrefs > 1) {
        const float t2 = c2.total * lf;
        if (t2 < best) {
            mode = DME_REF2_ONLY;
            best = t2;
        }
        dme_cost bc;
        bc.mvcost = c1.mvcost + c2.mvcost;
        bc.SAD = (float)bi_sad;
        bc.total = bc.SAD + lambda * bc.mvcost;
        bc.total *= lf;
        if (bc.total < best) {
            mode = DME_REF1AND2;
            best = bc.total;
        }
    }
    return best;
}

// Levels 1 and 0 of every superblock at once (DoSBDecn and DoLevelDecn,
// me_mode_decn.cpp:179-276) with DoME's predictor from the pre-merge level-2
// vectors (fv2).  One wave per superblock; the level-2 unit costs come from
// k_mode_l2.  Writes the split, the superblock cost, the merged vectors and
// modes, and the chosen units' SADs (fsb) for k_fast_eval.
// Level-1 results of a superblock, per wave in LDS.
struct FastUnits {
    dme_mv v[4][2];
    dme_cost c[4][2];
    int sad[4][2], bi[4], mode[4];
};

__global__ __launch_bounds__(256) void k_fast_mode(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                   int ntasks, int total, int pd, int prows) {
    extern __shared__ uint32_t s_fm[];
    __shared__ FastUnits s_units[4];
    const int wv = wave_id(), lane = threadIdx.x & (WAVE - 1);
    const int gi = xcd_block() * 4 + wv;
    if (gi >= total) return;
    uint32_t *s_pic = s_fm + wv * pd * prows;
    FastUnits &U = s_units[wv];
    const Task tk = tasks[find_task_wave(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    const int s = gi - tk.offset, XS = J.xsb, xs = s % XS, ys = s / XS;
    const int X = J.xnb, nr = J.nrefs;
    // level 2: the sum of the blocks' unit costs in raster order
    float sbc = 0.0f;
    for (int j = 4 * ys; j < 4 * ys + 4; ++j)
        for (int i = 4 * xs; i < 4 * xs + 4; ++i) sbc += J.ucost[(size_t)j * X + i];
    int split = 2;
    // level 1
    float c1 = 0.0f;
    for (int q = 0; q < 4; ++q) {
        const int i = 2 * xs + (q & 1), j = 2 * ys + (q >> 1), xb = i << 1, yb = j << 1;
        const Blk b = block_limits(J.blen[1], J.bsep[1], J.boff[1], J.W, J.H, i, j);
        const float lambda = J.lam1[(size_t)j * (X >> 1) + i];
        UnitLanes L = {};
        if (!b.empty()) {
            L = unit_lanes(b);
            WAVE_SYNC();
            stage_unit_pic(J, b, L, s_pic, pd);
        }
        UnitBest ub[2];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            if (r >= nr) {
                ub[r] = ub[0];
                continue;
            }
            const dme_mv *__restrict__ V2 = J.fv2[r];
            Cand4 g;
            g.c[0] = V2[(size_t)yb * X + xb];
            g.c[1] = V2[(size_t)yb * X + xb + 1];
            g.c[2] = V2[(size_t)(yb + 1) * X + xb];
            g.c[3] = V2[(size_t)(yb + 1) * X + xb + 1];
            ub[r] = unit_best(J, r, b, L, s_pic, pd, g, mode_pred(V2, X, xb, yb), lambda);
        }
        const int bs = (nr > 1 && !b.empty()) ? up8_bi_sad_s(J, b, L, s_pic, pd, ub[0].v, ub[1].v) : 0;
        int mode;
        c1 += unit_decn(nr, J.lf[1], lambda, ub[0].c, ub[1].c, bs, mode);
        if (lane == 0) {
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                U.v[q][r] = ub[r].v;
                U.c[q][r] = ub[r].c;
                U.sad[q][r] = ub[r].sad;
            }
            U.bi[q] = bs;
            U.mode[q] = mode;
        }
    }
    WAVE_SYNC();
    if (c1 <= sbc) {
        sbc = c1;
        split = 1;
    }
    // level 0: the guides are the level-1 results
    UnitBest u0[2];
    int m0, bs0;
    float c0;
    {
        const int xb = 4 * xs, yb = 4 * ys;
        const Blk b = block_limits(J.blen[0], J.bsep[0], J.boff[0], J.W, J.H, xs, ys);
        const float lambda = J.lam0[(size_t)ys * (X >> 2) + xs];
        UnitLanes L = {};
        if (!b.empty()) {
            L = unit_lanes(b);
            WAVE_SYNC();
            stage_unit_pic(J, b, L, s_pic, pd);
        }
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            if (r >= nr) {
                u0[r] = u0[0];
                continue;
            }
            Cand4 g;
#pragma unroll
            for (int q = 0; q < 4; ++q) g.c[q] = U.v[q][r];
            u0[r] = unit_best(J, r, b, L, s_pic, pd, g, mode_pred(J.fv2[r], X, xb, yb), lambda);
        }
        bs0 = (nr > 1 && !b.empty()) ? up8_bi_sad_s(J, b, L, s_pic, pd, u0[0].v, u0[1].v) : 0;
        c0 = unit_decn(nr, J.lf[0], lambda, u0[0].c, u0[1].c, bs0, m0);
    }
    if (c0 <= sbc) {
        sbc = c0;
        split = 0;
    }
    // merges into the level-2 arrays (DoLevelDecn's copy; the level-1/0 DC
    // is zero) and the chosen units' SADs
    if (split < 2) {
        // (values read from LDS first, then selected: no pointer selects
        // between LDS and registers)
        const dme_mv z0 = u0[0].v, z1 = u0[1].v;
        const int zs0 = u0[0].sad, zs1 = u0[1].sad;
        for (int t = lane; t < 16; t += WAVE) {
            const int bi = 4 * xs + (t & 3), bj = 4 * ys + (t >> 2);
            const size_t d = (size_t)bj * X + bi;
            const int q = ((t >> 3) << 1) | ((t & 3) >> 1);
            const dme_mv l0 = U.v[q][0], l1 = U.v[q][1];
            const int lm = U.mode[q];
            J.mode[d] = split == 0 ? m0 : lm;
            J.dc[0][d] = 0;
            J.vec[0][d] = split == 0 ? z0 : l0;
            if (nr > 1) J.vec[1][d] = split == 0 ? z1 : l1;
        }
        if (lane < 4 && (split == 1 || lane == 0)) {
            int *f = J.fsb + (size_t)s * kFsbInts + 3 * lane;
            const int ls0 = U.sad[lane][0], ls1 = U.sad[lane][1], lb = U.bi[lane];
            f[0] = split == 0 ? zs0 : ls0;
            f[1] = nr < 2 ? 0 : (split == 0 ? zs1 : ls1);
            f[2] = split == 0 ? bs0 : lb;
        }
    }
    if (lane == 0) {
        J.split[s] = split;
        J.sbc[s] = sbc;
    }
}

// The exact cost of every superblock split into units: each unit's costs
// with the predictor DoME forms (me_mode_decn.cpp:308-321) from the level-2
// vectors as DoModeDecn's raster pass sees them -- the final vectors of
// earlier superblocks, the pre-merge ones inside this one -- the reference's
// float semantics (one rounding, or the per-sample chain where its
// bounds-checked path adds per sample), DoUnitDecn and the sum in raster
// order.  One wave per superblock.
__global__ __launch_bounds__(256) void k_fast_eval(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                   int ntasks, int total) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int gi = xcd_block() * 4 + wave_id();
    if (gi >= total) return;
    const Task tk = tasks[find_task_wave(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    const int s = gi - tk.offset, XS = J.xsb, xs = s % XS, ys = s / XS;
    const int lev = J.split[s];
    if (lev == 2) return;
    const int X = J.xnb, nr = J.nrefs, UW = 2 * J.W, UH = 2 * J.H;
    const float lf = J.lf[lev];
    auto field = [&](int r, int u, int v) -> dme_mv {
        const int sx = u >> 2, sy = v >> 2;
        const bool earlier = sy < ys || (sy == ys && sx < xs);
        return earlier ? J.vec[r][(size_t)v * X + u] : J.fv2[r][(size_t)v * X + u];
    };
    float sbc = 0.0f;
    const int n = 1 << lev;  // units across
    for (int q = 0; q < n * n; ++q) {
        const int i = (xs << lev) + (q % n), j = (ys << lev) + (q / n);
        const int xb = i << (2 - lev), yb = j << (2 - lev);
        const Blk b = block_limits(J.blen[lev], J.bsep[lev], J.boff[lev], J.W, J.H, i, j);
        const float lambda = lev ? J.lam1[(size_t)j * (X >> 1) + i] : J.lam0[(size_t)j * (X >> 2) + i];
        const int *f = J.fsb + (size_t)s * kFsbInts + 3 * q;
        dme_cost c[2];
        c[1].SAD = c[1].mvcost = 0.0f;
        c[1].total = 100000000.0f;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            if (r >= nr) break;
            dme_mv pred;
            if (xb > 0 && yb > 0) pred = med3(field(r, xb - 1, yb), field(r, xb - 1, yb - 1), field(r, xb, yb - 1));
            else if (xb == 0 && yb > 0) pred = mean2(field(r, xb, yb - 1), field(r, xb + 1, yb - 1));
            else if (xb > 0 && yb == 0) pred = mean2(field(r, xb - 1, yb), field(r, xb - 1, yb + 1));
            else { pred.x = 0; pred.y = 0; }
            const dme_mv v = J.vec[r][(size_t)yb * X + xb];
            c[r].SAD = c[r].mvcost = 0.0f;
            c[r].total = 100000000.0f;
            if (!b.empty()) {
                const float mvcost = var_up(pred, v);
                const float start = mvcost * lambda;
                const UpStart us = up_start(b, v, J.step, UW, UH);
                const float t = us.bounds ? up8_chain(start, &J, r, b, us)
                                          : (float)f[r] + start;
                if (t < c[r].total) {
                    c[r].total = t;
                    c[r].mvcost = mvcost;
                    c[r].SAD = t - start;
                }
            }
        }
        int mode;
        sbc += unit_decn(nr, lf, lambda, c[0], c[nr > 1 ? 1 : 0], f[2], mode);
        const int ub = 1 << (2 - lev);  // the unit's blocks across
        for (int t = lane; t < ub * ub; t += WAVE) J.mode[(size_t)(yb + t / ub) * X + xb + t % ub] = mode;
    }
    if (lane == 0) J.sbc[s] = sbc;
}

// The MEData arrays of a batch packed into one buffer, one workgroup per
// array: 16-byte copies (device allocations are 256-aligned, offsets 16),
// then the 2-byte tail.
__global__ __launch_bounds__(256) void k_pack(const PackSeg *__restrict__ segs, uint8_t *__restrict__ dst) {
    const PackSeg sg = segs[blockIdx.x];
    const uint4 *__restrict__ s = reinterpret_cast<const uint4 *>(sg.src);
    uint4 *__restrict__ d = reinterpret_cast<uint4 *>(dst + sg.off);
    const unsigned long long n16 = sg.bytes / 16;
    for (unsigned long long i = threadIdx.x; i < n16; i += 256) d[i] = s[i];
    const unsigned long long t = (sg.bytes - 16 * n16) / 2;
    if (threadIdx.x < t)
        reinterpret_cast<uint16_t *>(dst + sg.off + 16 * n16)[threadIdx.x] =
            reinterpret_cast<const uint16_t *>(reinterpret_cast<const uint8_t *>(sg.src) + 16 * n16)[threadIdx.x];
}

// PictureCompressor::IntraModeAnalyse (picture_compress.cpp:300-320): the
// INTRA blocks of each job's final Mode array, one workgroup per job.
__global__ __launch_bounds__(256) void k_intra_count(const JobDesc *__restrict__ jobs) {
    __shared__ int s_n;
    const JobDesc &J = jobs[blockIdx.x];
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int n = J.xnb * J.ynb;
    int c = 0;
    for (int k = threadIdx.x; k < n; k += 256) c += J.mode[k] == DME_INTRA ? 1 : 0;
    c = wave_sum(c);
    if ((threadIdx.x & (WAVE - 1)) == 0) atomicAdd(&s_n, c);
    __syncthreads();
    if (threadIdx.x == 0) *J.icount = s_n;
}

hipError_t launch_intra_count(const JobDesc *jobs, int njobs, hipStream_t st) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_intra_count, dim3(njobs), dim3(256), 0, st, jobs);
    return hipGetLastError();
}

hipError_t launch_pack(const PackSeg *segs, int nseg, uint8_t *dst, hipStream_t st) {
    if (nseg <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack, dim3(nseg), dim3(256), 0, st, segs, dst);
    return hipGetLastError();
}

hipError_t launch_downconvert(const PlaneTask *pt, int npt, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_downconvert, dim3(total), dim3(256), 0, st, pt, npt, total);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Motion compensation: MotionCompensator::CompensatePicture and
// CompensateComponent (mot_comp.cpp:131-539) as a gather.  The reference walks
// the blocks of a row, scatters each block's weighted prediction into a strip
// of rows and flushes the strip; a sample's final sum is the sum over the units
// whose OBMC windows cover it -- at most two across (the overlap never exceeds
// a separation) by two down.  So each thread finds those units for its sample
// and sums their contributions directly: every sample is read and written
// once, no strip, no atomics.  The reference sums in 16-bit lanes that wrap;
// wrapping addition is order-independent, so an int sum truncated to 16 bits
// at the end is exact.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int mc_bchk(int v, int max) { return v < 0 ? 0 : (v >= max ? max - 1 : v); }

// Split mode as the reference's switch reads it: 0 and 1, anything else 2.
__device__ __forceinline__ int mc_split(int s) { return (s == 0 || s == 1) ? s : 2; }

__device__ __forceinline__ int mc_comp_of(const McComp *__restrict__ cs, int ncs, int tile) {
    int lo = 0, hi = ncs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cs[mid].offset <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One thread owns a column segment of MC_ROWS samples; a workgroup is 64
// columns by 4 x MC_ROWS rows.  The units covering the thread's column are
// found once per block row the segment meets (their weights across, modes,
// vectors and clamped columns do not change down the column), then each row
// only adds its taps.  Row tests and the weights down depend on the row
// alone, so they are uniform across the wave.
constexpr int MC_ROWS = kMcSegRows;

// Where a unit's prediction from one reference reads: the clamped columns j
// and x1 of the up-converted plane, the row offset, the bilinear fractions.
struct McTap {
    int j, x1, dy;
    int w00, w01, w10, w11;  // bilinear weights of (row, j), (row, x1), (row + 1, j), (row + 1, x1)
};

__device__ __forceinline__ McTap mc_tap(int prec, int x, dme_mv mv, int uw) {
    McTap t;
    int dx;
    t.w00 = 1;
    t.w01 = t.w10 = t.w11 = 0;
    if (prec == 0) {
        dx = 2 * mv.x;
        t.dy = 2 * mv.y;
    } else if (prec == 1) {
        dx = mv.x;
        t.dy = mv.y;
    } else {
        const int sh = prec - 1, one = 1 << sh;
        dx = mv.x >> sh;
        t.dy = mv.y >> sh;
        const int fx = mv.x & (one - 1), fy = mv.y & (one - 1);
        t.w00 = (one - fx) * (one - fy);
        t.w01 = fx * (one - fy);
        t.w10 = (one - fx) * fy;
        t.w11 = fx * fy;
    }
    // BChk against 2 size - 1 (mot_comp.cpp:1007-1008): columns 0 .. uw - 2
    t.j = min(max(2 * x + dx, 0), uw - 2);
    t.x1 = min(max(2 * x + dx + 1, 0), uw - 2);
    // x1 is j + 1 unless a clamp made it j: then column j takes both weights,
    // so that the taps always read columns j and j + 1 (j + 1 <= uw - 1)
    if (t.x1 == t.j) {
        t.w00 += t.w01;
        t.w10 += t.w11;
        t.w01 = t.w11 = 0;
    }
    return t;
}

// BlockPixelPred (mot_comp.cpp:846-1230) for one sample of row y; BIL:
// quarter or eighth pel (bilinear between four up-converted samples), else
// one sample.  Rows clamp like columns (BChk), so any y reads inside the plane.
// The plane's bytes are samples biased by 128: the bilinear weights sum to
// 2^(2 sh), so the bias passes through the rounded shift whole.  up is the
// same for the whole wave and the offsets are unsigned 32-bit, so each load
// adds a lane offset to a scalar base.
template <bool BIL>
__device__ __forceinline__ int mc_eval(const uint8_t DME_G *__restrict__ up, int uw, int ty, int sh, int y,
                                       const McTap &t) {
    const unsigned o0 = (unsigned)(mc_bchk(2 * y + t.dy, ty) * uw) + (unsigned)t.j;
    if (!BIL) return (int)up[o0] - 128;
    const unsigned o1 = (unsigned)(mc_bchk(2 * y + t.dy + 1, ty) * uw) + (unsigned)t.j;
    const int v = t.w00 * up[o0] + t.w01 * up[o0 + 1u] + t.w10 * up[o1] + t.w11 * up[o1 + 1u];
    return ((v + (1 << (2 * sh - 1))) >> (2 * sh)) - 128;
}

// Adds one unit's weighted prediction to the segment's rows; rows outside the
// unit's window have weight 0 and are skipped (a uniform branch).  Lanes
// predict from reference 1 (use1, tap t1, plane up1) and / or reference 2
// (use2, t2, up2), both planes the wave's own; bi-prediction averages the two.
template <bool BIL>
__device__ __forceinline__ void mc_rows(int (&S)[kMcSegRows], const uint8_t DME_G *__restrict__ up1,
                                        const uint8_t DME_G *__restrict__ up2, int uw, int ty, int sh, int y0,
                                        const int (&W)[kMcSegRows], int H, const McTap &t1, const McTap &t2, bool use1,
                                        bool use2) {
#pragma unroll
    for (int i = 0; i < kMcSegRows; ++i) {
        const int y = y0 + i;
        if (W[i] == 0) continue;
        int v1 = 0, v2 = 0;
        if (use1) v1 = mc_eval<BIL>(up1, uw, ty, sh, y, t1);
        if (use2) v2 = mc_eval<BIL>(up2, uw, ty, sh, y, t2);
        const int val = (use1 && use2) ? (v1 + v2 + 1) >> 1 : v1 + v2;
        S[i] += val * (W[i] * H);
    }
}

__global__ __launch_bounds__(256) void k_mc(const McComp *__restrict__ cs, int ncs, int total) {
    const int tile = blockIdx.x;
    if (tile >= total) return;
    const McComp &C = cs[mc_comp_of(cs, ncs, tile)];
    const int local = tile - C.offset;
    const int x = (local % C.tiles_x) * 64 + (threadIdx.x & 63);
    const int y0 = (local / C.tiles_x) * (4 * MC_ROWS) + wave_id() * MC_ROWS;
    if (x >= C.lx || y0 >= C.ly) return;
    int16_t DME_G *col = C.out + x;
    if (x >= C.pw || y0 >= C.ph) {
        // subtracting zeroes the padding (mot_comp.cpp:449-455, 511-522);
        // adding replicates the last column and row, in k_mc_pad
        if (C.direction == DME_MC_SUBTRACT)
#pragma unroll
            for (int i = 0; i < MC_ROWS; ++i)
                if (y0 + i < C.ly) col[(size_t)(y0 + i) * C.lx] = 0;
        return;
    }
    const int uw = 2 * C.pw, ty = 2 * C.ph - 1;
    const int yend = min(y0 + MC_ROWS, C.ph);  // rows of the picture in this segment
    const int c = (x + C.xoff) / C.xbsep;      // the block column whose separation holds x
    int S[MC_ROWS];
#pragma unroll
    for (int i = 0; i < MC_ROWS; ++i) S[i] = 0;
    const int rb0 = max(0, (y0 + C.yoff) / C.ybsep - 1);
    const int rb1 = min(C.nby - 1, (yend - 1 + C.yoff) / C.ybsep);
    for (int rb = rb0; rb <= rb1; ++rb) {
        const int py = rb * C.ybsep - C.yoff;
        const int ylo = max(py, y0), yhi = min(py + C.yblen, yend);
        if (ylo >= yhi) continue;
        // weight class down (mot_comp.cpp:339-371): top, inside, bottom
        const int ycls = py < 0 ? 0 : (rb + 1 < C.nby ? 1 : 2);
        // the weights down of the segment's rows (0 outside the window)
        const int DME_G *V = C.vw + ycls * C.yblen;
        int W[MC_ROWS];
#pragma unroll
        for (int i = 0; i < MC_ROWS; ++i) {
            const int y = y0 + i;
            W[i] = (y >= ylo && y < yhi) ? V[min(max(y - py, 0), C.yblen - 1)] : 0;
        }
        const int DME_G *srow = C.split + (size_t)(rb >> 2) * C.xsb;
        // units across: the one holding block c (if in use), then the one
        // before, in this superblock or the previous one (both splits are
        // loaded at once)
        const int cb = min(c, C.nbx - 1);
        const int sa = mc_split(srow[cb >> 2]), sbp = mc_split(srow[max((cb >> 2) - 1, 0)]);
        int ub[2] = {-1, -1}, sp[2] = {0, 0}, pb;
        if (c < C.nbx) {
            const int u1 = c & ~((4 >> sa) - 1);
            ub[1] = u1;
            sp[1] = sa;
            pb = u1 - 1;
        } else {
            pb = C.nbx - 1;
        }
        if (pb >= 0) {
            const int s0 = (pb >> 2) == (cb >> 2) ? sa : sbp;
            const int k0 = 4 >> s0, u0 = pb & ~(k0 - 1);
            if (x - (u0 * C.xbsep - C.xoff) < k0 * C.xbsep + C.xov) {
                ub[0] = u0;
                sp[0] = s0;
            }
        }
        for (int q = 0; q < 2; ++q) {
            const int u = ub[q];
            if (u < 0) continue;
            const int s = sp[q], ku = 4 >> s;
            const int px = u * C.xbsep - C.xoff;
            const int xcls = (px >= 0 && u + ku < C.nbx) ? 1 : (px < 0 ? 0 : 2);
            const int H = C.hw[(s * 3 + xcls) * C.hpitch + (x - px)];
            const size_t b = (size_t)rb * C.xnb + u;
            const int4 rv = C.unit[2 * b], rm = C.unit[2 * b + 1];
            const int mode = rm.x;
            if (mode == DME_REF1_ONLY || mode == DME_REF2_ONLY || mode == DME_REF1AND2) {
                dme_mv m1, m2;
                m1.x = rv.x >> C.shift;
                m1.y = rv.y >> C.shift;
                m2.x = rv.z >> C.shift;
                m2.y = rv.w >> C.shift;
                const McTap t1 = mc_tap(C.prec, x, m1, uw), t2 = mc_tap(C.prec, x, m2, uw);
                const bool use1 = mode != DME_REF2_ONLY, use2 = mode != DME_REF1_ONLY;
                const int sh = C.prec - 1;
                if (C.prec >= 2)
                    mc_rows<true>(S, C.up[0], C.up[1], uw, ty, sh, y0, W, H, t1, t2, use1, use2);
                else
                    mc_rows<false>(S, C.up[0], C.up[1], uw, ty, sh, y0, W, H, t1, t2, use1, use2);
            } else {
                const int dcv = C.comp == 0 ? rm.y : (C.comp == 1 ? rm.z : rm.w);  // DCBlock
#pragma unroll
                for (int i = 0; i < MC_ROWS; ++i) S[i] += dcv * (W[i] * H);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < MC_ROWS; ++i) {
        const int y = y0 + i;
        if (y >= C.ly) break;
        int16_t DME_G *o = col + (size_t)y * C.lx;
        if (y >= C.ph) {
            if (C.direction == DME_MC_SUBTRACT) *o = 0;
            continue;
        }
        const int s16 = (int16_t)S[i];
        if (C.direction == DME_MC_SUBTRACT) {
            *o = (int16_t)(*o - (int16_t)((s16 + 32) >> 6));
        } else {
            // the MMX add path rounds in 16-bit lanes over the first
            // (width & ~3) samples (mot_comp_mmx.cpp:408-424)
            const int p = (C.mmx && x < (C.pw & ~3)) ? ((int16_t)(s16 + 32)) >> 6 : (s16 + 32) >> 6;
            *o = (int16_t)(*o + (int16_t)p);
        }
    }
}

// Packs the MvData k_mc reads per block into two 16-byte records, so that a
// unit's vectors, mode and DC values arrive with one round trip.
__global__ __launch_bounds__(256) void k_mc_pack(const McPackJob *__restrict__ jobs, int njobs, int total) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].offset <= i) lo = mid;
        else hi = mid - 1;
    }
    const McPackJob &J = jobs[lo];
    const int b = i - J.offset;
    const dme_mv a = J.vec[0][b], v = J.vec[1][b];
    J.out[2 * b] = make_int4(a.x, a.y, v.x, v.y);
    J.out[2 * b + 1] = make_int4(J.mode[b], J.dc[0][b], J.dc[1][b], J.dc[2][b]);
}

// Adding: samples beyond the picture take the last column's value, rows below
// it the last row's (mot_comp.cpp:484-488, 523-537).  Runs after k_mc.
__global__ __launch_bounds__(256) void k_mc_pad(const McComp *__restrict__ cs, int ncs, int total) {
    const int tile = blockIdx.x;
    if (tile >= total) return;
    const McComp &C = cs[mc_comp_of(cs, ncs, tile)];
    const int local = tile - C.offset;
    const int x = (local % C.tiles_x) * 64 + (threadIdx.x & 63);
    const int y = (local / C.tiles_x) * 4 + wave_id();
    if (x >= C.lx || y >= C.ly || (x < C.pw && y < C.ph)) return;
    C.out[(size_t)y * C.lx + x] = C.out[(size_t)min(y, C.ph - 1) * C.lx + min(x, C.pw - 1)];
}

hipError_t launch_pel_full(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int pp, int rp, int rows,
                           hipStream_t st) {
    if (total <= 0) return hipSuccess;
    FullLds ly;
    ly.pp = pp;
    ly.rp = rp;
    ly.rows = rows;
    ly.pic = pp * 64;  // blocks of up to 64 rows
    const size_t bytes = sizeof(int) * (8 + 26) + sizeof(uint32_t) * ((size_t)ly.pic + (size_t)rows * rp);
    if (bytes > 160 * 1024) return hipErrorInvalidValue;
    static size_t attr = 0;  // the largest size set so far
    if (bytes > attr) {
        hipError_t e = hipFuncSetAttribute((const void *)k_pel_full, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)bytes);
        if (e != hipSuccess) return e;
        attr = bytes;
    }
    hipLaunchKernelGGL(k_pel_full, dim3(total), dim3(256), bytes, st, jobs, tasks, ntasks, total, ly);
    return hipGetLastError();
}

hipError_t launch_mc_pack(const McPackJob *jobs, int njobs, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mc_pack, dim3((total + 255) / 256), dim3(256), 0, st, jobs, njobs, total);
    return hipGetLastError();
}

hipError_t launch_mc(const McComp *cs, int ncs, int total, bool pad, hipStream_t st) {
    // total counts 64 x 4 tiles for k_mc_pad and 64 x (4 MC_ROWS) tiles for k_mc
    if (total <= 0) return hipSuccess;
    if (pad) hipLaunchKernelGGL(k_mc_pad, dim3(total), dim3(256), 0, st, cs, ncs, total);
    else hipLaunchKernelGGL(k_mc, dim3(total), dim3(256), 0, st, cs, ncs, total);
    return hipGetLastError();
}

hipError_t launch_upconvert(const PlaneTask *pt, int npt, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_upconvert<false>, dim3(total), dim3(256), 0, st, pt, npt, total);
    return hipGetLastError();
}

// The byte planes of motion compensation (PlaneTask::out holds 4 W H bytes).
hipError_t launch_upconvert8(const PlaneTask *pt, int npt, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_upconvert<true>, dim3(total), dim3(256), 0, st, pt, npt, total);
    return hipGetLastError();
}

hipError_t launch_lambda(const JobDesc *jobs, const Task *tasks, int ntasks, int *flags, hipStream_t st) {
    if (ntasks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lambda, dim3(ntasks), dim3(1024), 0, st, jobs, tasks, flags);
    return hipGetLastError();
}

hipError_t launch_bytes(const PlaneTask *pt, int npt, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bytes, dim3(total), dim3(256), 0, st, pt, npt, total);
    return hipGetLastError();
}

// Dynamic LDS of k_pel_bulk (four waves) for blocks up to bx x by at
// `level`, with its layout.
static size_t pel_bulk_lds(int level, int max_blen_x, int max_blen_y, bool has_guide, bool u8, PelLds &ly) {
    const int rad = min(level + 1, 5), R = rad + 1, D = 2 * R + 1;
    const int ntab = has_guide ? 2 : 1;
    // dwords of a packed row: per = samples per dword; the SAD reads 4-dword
    // chunks of the block row and one dword more of the region row, from
    // dword offset (2R) / per
    const int per = u8 ? 4 : 2;
    const int nch = ((max_blen_x + per - 1) / per + 3) / 4;
    ly.pp = 4 * nch;
    ly.rp = max((max_blen_x + 2 * R + per - 1) / per, (2 * R) / per + 4 * nch + 1) | 1;
    ly.pic = ly.pp * max_blen_y;
    ly.reg = ly.rp * (max_blen_y + 2 * R);
    ly.keys = 2 * ly.reg;
    ly.tab = ntab * D * D;
    return sizeof(uint32_t) * 4 * (size_t)(ly.pic + ly.keys + ly.tab);
}
constexpr size_t kPelBulkLdsMax = 64 * 1024;

// Dynamic LDS of k_pel_bulk8 (four waves) for blocks up to bx x by.  A
// lane reads region dwords q .. q + nd, q <= 1; the staging writes up to ncm.
static size_t pel_bulk8_lds(int max_blen_x, int max_blen_y, Pb8Lds &ly) {
    const int nd = (max_blen_x + 3) / 4;
    ly.ndm = nd;
    ly.ncm = (max_blen_x + 4 + 3 + 3) / 4;  // region of xl + 4 samples from a shift up to 3
    ly.pp = (nd + 3) & ~3;
    ly.prows = max_blen_y;
    ly.rp = std::max(ly.ncm, nd + 2) | 1;
    ly.rrows = max_blen_y + 4;
    const int item = (ly.pp * ly.prows + 2 * ly.rp * ly.rrows + 50 + 18 + 3) & ~3;
    return sizeof(uint32_t) * 4 * (size_t)PB_BLOCKS * item;
}
constexpr size_t kPelBulk8LdsMax = 159 * 1024;

static hipError_t launch_pel_bulk8(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int max_blen_x,
                                   int max_blen_y, hipStream_t st) {
    Pb8Lds ly;
    const size_t lds = pel_bulk8_lds(max_blen_x, max_blen_y, ly);
    if (lds > kPelBulk8LdsMax) return hipErrorInvalidValue;
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)k_pel_bulk8, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           159 * 1024);
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL(k_pel_bulk8, dim3((total + 4 * PB_BLOCKS - 1) / (4 * PB_BLOCKS)), dim3(256), lds, st, jobs,
                       tasks, ntasks, total, ly);
    return hipGetLastError();
}

// u8: level 0 of luma jobs (the pictures themselves, whose samples were
// checked to 8-bit range at registration), from the byte planes; else 16-bit
// samples (the pyramid levels, and every level of combined-component jobs).
hipError_t launch_pel_bulk(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int level, int max_blen_x,
                           int max_blen_y, bool has_guide, bool u8, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    if (max_blen_x < 1 || max_blen_x > PEL_MAXB || max_blen_y < 1 || max_blen_y > PEL_MAXB) return hipErrorInvalidValue;
    if (u8 && level != 0) return hipErrorInvalidValue;
    PelLds ly;
    const size_t lds = pel_bulk_lds(level, max_blen_x, max_blen_y, has_guide, u8, ly);
    if (lds > kPelBulkLdsMax) return hipErrorInvalidValue;
    const dim3 grid((total + 4 * PW_PEL_BULK - 1) / (4 * PW_PEL_BULK));
    if (u8) return launch_pel_bulk8(jobs, tasks, ntasks, total, max_blen_x, max_blen_y, st);
    hipLaunchKernelGGL(k_pel_bulk<false>, grid, dim3(256), lds, st, jobs, tasks, ntasks, total, level, ly);
    return hipGetLastError();
}

// LDS of k_pel_wave: a fixed part for max_rows block rows and a part per
// queued miss (at least one must fit).
static void pel_wave_lds(int level, int max_rows, int max_blen_x, int max_blen_y, int &stage_elems, size_t &fixed,
                         size_t &per_miss) {
    const int rad = min(level + 1, 5), n = (2 * rad + 1) * (2 * rad + 1);
    stage_elems = (max_blen_x * max_blen_y + (max_blen_x + 2 * rad) * (max_blen_y + 2 * rad) + 1) & ~1;
    fixed = sizeof(int) * (6 * (size_t)max_rows + 2 * (size_t)max_rows * PEL_REC);
    per_miss = sizeof(PelMiss) + sizeof(int) * n + sizeof(int16_t) * stage_elems;
}
constexpr size_t kPelWaveLdsMax = 160 * 1024 - 1024;  // static LDS of the kernel stays below 1 KiB

hipError_t launch_pel_wave(const JobDesc *jobs, const Task *tasks, int ntasks, int level, int max_rows, int max_blen_x,
                           int max_blen_y, int threads, hipStream_t st) {
    if (ntasks <= 0) return hipSuccess;
    if (threads < 64 || threads > 1024 || threads % 64) return hipErrorInvalidValue;  // the kernel's launch bound
    int stage_elems = 0;
    size_t fixed = 0, per_miss = 0;
    pel_wave_lds(level, max_rows, max_blen_x, max_blen_y, stage_elems, fixed, per_miss);
    const size_t budget = kPelWaveLdsMax;
    if (fixed + per_miss > budget) return hipErrorInvalidValue;
    const int miss_cap = (int)min((size_t)256, (budget - fixed) / per_miss);
    const size_t lds = fixed + per_miss * miss_cap;
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)k_pel_wave, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)budget);
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL(k_pel_wave, dim3(ntasks), dim3(threads), lds, st, jobs, tasks, level, max_rows, miss_cap,
                       stage_elems);
    return hipGetLastError();
}

constexpr size_t kTableLdsMax = 159 * 1024;  // table and sub-pel wavefront kernels (their attribute)

// LDS of k_subpel_bulk (four waves) for blocks up to bx x by: a phase region
// spans the block plus two positions per axis; the SAD reads 4-dword chunks
// of a picture row and one dword more of a region row.
static size_t subpel_bulk_lds(int max_blen_x, int max_blen_y, int &pp, int &prows, int &qrows, int &qd) {
    const int nch = ((max_blen_x + 3) / 4 + 3) / 4;
    pp = 4 * nch;
    prows = max_blen_y;
    qrows = max_blen_y + 2;
    qd = 4 * nch + 1;
    return sizeof(uint32_t) * 4 * ((size_t)pp * prows + 16 * (size_t)qrows * qd);
}

static size_t subpel_wave_lds(int max_rows) {
    return sizeof(int) * (8 + 2 * SUB_SLOT) * (size_t)max_rows + sizeof(SubJob) * SUB_QCAP;
}

hipError_t launch_subpel_bulk(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int max_blen_x,
                              int max_blen_y, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    if (max_blen_x < 1 || max_blen_x > PEL_MAXB || max_blen_y < 1 || max_blen_y > PEL_MAXB) return hipErrorInvalidValue;
    int pp, prows, qrows, qd;
    const size_t lds = subpel_bulk_lds(max_blen_x, max_blen_y, pp, prows, qrows, qd);
    if (lds > kTableLdsMax) return hipErrorInvalidValue;
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)k_subpel_bulk, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           159 * 1024);
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL(k_subpel_bulk, dim3((total + 3) / 4), dim3(256), lds, st, jobs, tasks, ntasks, total, pp, prows,
                       qrows, qd);
    return hipGetLastError();
}

hipError_t launch_subpel_table16(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int max_blen_x,
                                 int max_blen_y, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    if (max_blen_x < 1 || max_blen_x > PEL_MAXB || max_blen_y < 1 || max_blen_y > PEL_MAXB) return hipErrorInvalidValue;
    // per wave: the block and the window, int16, in whole dwords
    const int words = (((max_blen_x * max_blen_y + 1) & ~1) + subpel16_window(max_blen_x) * subpel16_window(max_blen_y) +
                       1) & ~1;
    const size_t lds = sizeof(int16_t) * 4 * (size_t)words;  // 4 waves: at most 2 * (1024 + 68 * 68) * 4 bytes
    hipLaunchKernelGGL(k_subpel_table16, dim3((total + 3) / 4), dim3(256), lds, st, jobs, tasks, ntasks, total, words);
    return hipGetLastError();
}

hipError_t launch_mode_tables16(const JobDesc *jobs, const Task *tasks, int ntasks, int total, bool eighth,
                                hipStream_t st) {
    if (total <= 0) return hipSuccess;
    if (eighth) hipLaunchKernelGGL(k_mode_tables16<true>, dim3((total + 3) / 4), dim3(256), 0, st, jobs, tasks, ntasks,
                                   total);
    else hipLaunchKernelGGL(k_mode_tables16<false>, dim3((total + 3) / 4), dim3(256), 0, st, jobs, tasks, ntasks, total);
    return hipGetLastError();
}

hipError_t launch_combine(const CombTask *ct, int nct, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_combine, dim3(total), dim3(256), 0, st, ct, nct, total);
    return hipGetLastError();
}

hipError_t launch_check_range(const int16_t *y, size_t n, int *flag, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const int blocks = (int)std::min<size_t>((n + 1023) / 1024, 4096);
    hipLaunchKernelGGL(k_check_range, dim3(blocks), dim3(256), 0, st, y, n, flag);
    return hipGetLastError();
}

hipError_t launch_build_qp(const PlaneTask *pt, int npt, int total, bool up8, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    if (up8) hipLaunchKernelGGL(k_build_qp<true>, dim3(total), dim3(256), 0, st, pt, npt, total);
    else hipLaunchKernelGGL(k_build_qp<false>, dim3(total), dim3(256), 0, st, pt, npt, total);
    return hipGetLastError();
}

hipError_t launch_subpel_wave(const JobDesc *jobs, const Task *tasks, int ntasks, int max_rows, int threads, int kinds,
                              hipStream_t st) {
    if (ntasks <= 0) return hipSuccess;
    if (threads < 64 || threads > 512 || threads % 64) return hipErrorInvalidValue;  // the kernel's launch bound
    const size_t lds = subpel_wave_lds(max_rows);
    if (lds > kTableLdsMax) return hipErrorInvalidValue;
    static bool attr = false;
    if (!attr) {
        for (const void *f : {(const void *)k_subpel_wave<false, false>, (const void *)k_subpel_wave<true, false>,
                              (const void *)k_subpel_wave<false, true>, (const void *)k_subpel_wave<true, true>}) {
            hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 159 * 1024);
            if (e != hipSuccess) return e;
        }
        attr = true;
    }
    if (kinds & DME_KIND_BELOW_EIGHTH)
        hipLaunchKernelGGL((k_subpel_wave<false, false>), dim3(ntasks), dim3(threads), lds, st, jobs, tasks, max_rows);
    if (kinds & DME_KIND_EIGHTH)
        hipLaunchKernelGGL((k_subpel_wave<true, false>), dim3(ntasks), dim3(threads), lds, st, jobs, tasks, max_rows);
    if (kinds & DME_KIND_COMB_BELOW_EIGHTH)
        hipLaunchKernelGGL((k_subpel_wave<false, true>), dim3(ntasks), dim3(threads), lds, st, jobs, tasks, max_rows);
    if (kinds & DME_KIND_COMB_EIGHTH)
        hipLaunchKernelGGL((k_subpel_wave<true, true>), dim3(ntasks), dim3(threads), lds, st, jobs, tasks, max_rows);
    return hipGetLastError();
}

hipError_t launch_shift_vectors(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int left,
                                hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_shift_vectors, dim3((total + 255) / 256), dim3(256), 0, st, jobs, tasks, ntasks, total, left);
    return hipGetLastError();
}

hipError_t launch_mode_l2(const JobDesc *jobs, const Task *tasks, int ntasks, int total, bool up8, int kinds,
                          hipStream_t st) {
    if (total <= 0) return hipSuccess;
    const dim3 grid((total + 4 * PW_MODE_L2 - 1) / (4 * PW_MODE_L2));
    if (up8) {  // the fast form: quarter or half pel only, on luma
        hipLaunchKernelGGL((k_mode_l2<true, false, false>), grid, dim3(256), 0, st, jobs, tasks, ntasks, total);
        return hipGetLastError();
    }
    if (kinds & DME_KIND_BELOW_EIGHTH)
        hipLaunchKernelGGL((k_mode_l2<false, false, false>), grid, dim3(256), 0, st, jobs, tasks, ntasks, total);
    if (kinds & DME_KIND_EIGHTH)
        hipLaunchKernelGGL((k_mode_l2<false, true, false>), grid, dim3(256), 0, st, jobs, tasks, ntasks, total);
    if (kinds & DME_KIND_COMB_BELOW_EIGHTH)
        hipLaunchKernelGGL((k_mode_l2<false, false, true>), grid, dim3(256), 0, st, jobs, tasks, ntasks, total);
    if (kinds & DME_KIND_COMB_EIGHTH)
        hipLaunchKernelGGL((k_mode_l2<false, true, true>), grid, dim3(256), 0, st, jobs, tasks, ntasks, total);
    return hipGetLastError();
}

// LDS of k_mode_tables (four waves) for level-0 units up to x0 x y0 and
// level-1 units up to x1 x y1.
static size_t mode_tables_lds(int max_x0, int max_y0, int max_x1, int max_y1, MtLds &ly) {
    ly.pd = (std::max(max_x0, max_x1) + 3) / 4;
    ly.prows = std::max(max_y0, max_y1);
    ly.spd = (max_x1 + 3) / 4;
    ly.slot = ly.spd * max_y1;
    ly.wave = ly.pd * ly.prows + 8 * ly.slot;
    return sizeof(uint32_t) * 4 * (size_t)ly.wave;
}

hipError_t launch_mode_tables(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int max_x0, int max_y0,
                              int max_x1, int max_y1, int kinds, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    if (max_x0 < 1 || max_y0 < 1 || max_x1 < 1 || max_y1 < 1 || max_x0 > 256 || max_x1 > 256) return hipErrorInvalidValue;
    MtLds ly;
    const size_t lds = mode_tables_lds(max_x0, max_y0, max_x1, max_y1, ly);
    if (lds > kTableLdsMax) return hipErrorInvalidValue;
    static bool attr = false;
    if (!attr) {
        for (const void *f : {(const void *)k_mode_tables<false>, (const void *)k_mode_tables<true>}) {
            hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, 159 * 1024);
            if (e != hipSuccess) return e;
        }
        attr = true;
    }
    const dim3 grid((total + 4 * PW_MODE_TABLES - 1) / (4 * PW_MODE_TABLES));
    if (kinds & DME_KIND_BELOW_EIGHTH)
        hipLaunchKernelGGL(k_mode_tables<false>, grid, dim3(256), lds, st, jobs, tasks, ntasks, total, ly);
    if (kinds & DME_KIND_EIGHTH)
        hipLaunchKernelGGL(k_mode_tables<true>, grid, dim3(256), lds, st, jobs, tasks, ntasks, total, ly);
    return hipGetLastError();
}

hipError_t launch_mode_wave(const JobDesc *jobs, int njobs, int threads, int kinds, hipStream_t st) {
    if (threads < 64 || threads > 1024 || threads % 64) return hipErrorInvalidValue;  // the kernel's launch bound
    if (kinds & DME_KIND_BELOW_EIGHTH)
        hipLaunchKernelGGL((k_mode_wave<false, false>), dim3(njobs), dim3(threads), 0, st, jobs);
    if (kinds & DME_KIND_EIGHTH) hipLaunchKernelGGL((k_mode_wave<true, false>), dim3(njobs), dim3(threads), 0, st, jobs);
    if (kinds & DME_KIND_COMB_BELOW_EIGHTH)
        hipLaunchKernelGGL((k_mode_wave<false, true>), dim3(njobs), dim3(threads), 0, st, jobs);
    if (kinds & DME_KIND_COMB_EIGHTH)
        hipLaunchKernelGGL((k_mode_wave<true, true>), dim3(njobs), dim3(threads), 0, st, jobs);
    return hipGetLastError();
}

hipError_t launch_block_dc(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int max_width,
                           hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_block_dc, dim3(total), dim3(256), (size_t)max_width * sizeof(int), st, jobs, tasks, ntasks,
                       total);
    return hipGetLastError();
}

hipError_t launch_dc_smooth(const JobDesc *jobs, const Task *tasks, int ntasks, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dc_smooth, dim3((total + 255) / 256), dim3(256), 0, st, jobs, tasks, ntasks, total);
    return hipGetLastError();
}

hipError_t launch_lp_features(const JobDesc *jobs, const Task *tasks, int ntasks, int total, uint16_t *feat, float *lab,
                              hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lp_features, dim3((total + 255) / 256), dim3(256), 0, st, jobs, tasks, ntasks, total, feat, lab);
    return hipGetLastError();
}

// LDS of k_lp_refine (four waves): the SAD reads 4-dword chunks of a
// picture row and, from dword offset (2 * radius) / 4, one dword more of a
// region row.
static size_t lp_refine_lds(int max_blen_x, int max_blen_y, LpLds &ly) {
    const int nd = (max_blen_x + 3) / 4;
    ly.pp = (nd + 3) & ~3;
    ly.prows = max_blen_y;
    // a lane reads dwords q .. q + nd of a box row, q <= (2 LP_BOX + 3) / 4
    ly.rp = ((2 * LP_BOX + 3) / 4 + nd + 1) | 1;
    ly.brows = max_blen_y + 2 * LP_BOX;
    ly.zp = (2 + nd) | 1;
    ly.zrows = max_blen_y + 4;
    const int item = (ly.pp * ly.prows + ly.rp * ly.brows + ly.zp * ly.zrows + 3) & ~3;
    return sizeof(uint32_t) * 4 * (size_t)(LP_ITEMS * item + 8);
}
constexpr size_t kLpRefineLdsMax = 159 * 1024;

// LDS of k_fast_mode (four waves): the packed picture rows of the largest
// unit of the launch (level-1 and level-0 lengths).
static size_t fast_mode_lds(int max_x, int max_y, int &pd) {
    pd = (max_x + 3) / 4;
    return sizeof(uint32_t) * 4 * (size_t)pd * max_y;
}
constexpr size_t kFastModeLdsMax = 60 * 1024;  // beside its static per-wave unit results

// LDS of k_fast_subpel (four waves) for blocks up to bx x by: 2 by + 3
// staged up rows of 2 ceil(bx / 4) + 3 dwords (the up samples of the block's
// candidates from the dword before them, and the two dwords quad_u8 reads
// past a lane's first sample), then the picture block.
static size_t fast_subpel_lds(int max_blen_x, int max_blen_y, FsLds &ly) {
    ly.rows = 2 * max_blen_y + 3;
    ly.pitch = 2 * ((max_blen_x + 3) / 4) + 3;
    ly.pd = (max_blen_x + 3) / 4;
    ly.wave = ly.rows * ly.pitch + ly.pd * max_blen_y;
    return sizeof(uint32_t) * 4 * (size_t)ly.wave;
}
constexpr size_t kFastSubpelLdsMax = 64 * 1024;

// The limits of one job's launches, checked on the host before anything is
// queued: "" if every kernel of the job's stages fits, else the limit.
std::string launch_limits(const Geometry &g, bool full_search, bool learned, bool combined) {
    const int bx = g.blen[2][0], by = g.blen[2][1];
    if (bx > PEL_MAXB || by > PEL_MAXB)
        return "block lengths above " + std::to_string(PEL_MAXB) + " after rationalising";
    PelLds pl;
    for (int l = full_search ? 0 : g.depth; l >= 0; --l) {
        if (pel_bulk_lds(l, bx, by, l < g.depth, l == 0 && !combined, pl) > kPelBulkLdsMax)
            return "pel tables of level " + std::to_string(l) + " exceed 64 KiB of LDS";
        int se = 0;
        size_t fixed = 0, per_miss = 0;
        pel_wave_lds(l, g.lyb[l], bx, by, se, fixed, per_miss);
        if (fixed + per_miss > kPelWaveLdsMax)
            return "pel wavefront of level " + std::to_string(l) + ": " + std::to_string(g.lyb[l]) +
                   " block rows exceed one workgroup's LDS";
        if (learned) break;  // levels 1 and 0: the learned refinement
    }
    int pp, prows, qrows, qd;
    if (subpel_bulk_lds(bx, by, pp, prows, qrows, qd) > kTableLdsMax) return "sub-pel tables exceed the LDS";
    if (subpel_wave_lds(g.ynb) > kTableLdsMax)
        return "sub-pel wavefront: " + std::to_string(g.ynb) + " block rows exceed one workgroup's LDS";
    MtLds ml;
    if (g.blen[0][0] > 256 || g.blen[1][0] > 256 ||
        mode_tables_lds(g.blen[0][0], g.blen[0][1], g.blen[1][0], g.blen[1][1], ml) > kTableLdsMax)
        return "mode-decision units exceed the LDS of the unit tables";
    LpLds ll;
    if (learned && lp_refine_lds(bx, by, ll) > kLpRefineLdsMax) return "learned refinement exceeds 159 KiB of LDS";
    int fpd;
    if (learned && fast_mode_lds(std::max(g.blen[0][0], g.blen[1][0]), std::max(g.blen[0][1], g.blen[1][1]), fpd) >
                       kFastModeLdsMax)
        return "the learned form's superblock units exceed its LDS";
    FsLds fl;
    if (learned && fast_subpel_lds(bx, by, fl) > kFastSubpelLdsMax)
        return "the learned form's sub-pel staging exceeds 64 KiB of LDS";
    if ((size_t)g.W * sizeof(int) > 64 * 1024) return "picture width exceeds the block-DC row buffer";
    return std::string();
}

hipError_t launch_lp_refine(const JobDesc *jobs, const Task *tasks, int ntasks, int total, const float *out4,
                            unsigned long long *stats, int max_blen_x, int max_blen_y, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    if (max_blen_x < 1 || max_blen_x > PEL_MAXB || max_blen_y < 1 || max_blen_y > PEL_MAXB) return hipErrorInvalidValue;
    LpLds ly;
    const size_t lds = lp_refine_lds(max_blen_x, max_blen_y, ly);
    if (lds > kLpRefineLdsMax) return hipErrorInvalidValue;
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)k_lp_refine, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           159 * 1024);
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL(k_lp_refine, dim3((total + 4 * LP_ITEMS - 1) / (4 * LP_ITEMS)), dim3(256), lds, st, jobs, tasks,
                       ntasks, total, out4, stats, ly);
    return hipGetLastError();
}

hipError_t launch_lp_pel_costs(const JobDesc *jobs, const Task *tasks, int ntasks, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lp_pel_costs, dim3((total + 255) / 256), dim3(256), 0, st, jobs, tasks, ntasks, total);
    return hipGetLastError();
}


// ---- the fast form of the learned path ----------------------------------------


hipError_t launch_fast_subpel(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int max_blen_x,
                              int max_blen_y, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    if (max_blen_x < 1 || max_blen_x > PEL_MAXB || max_blen_y < 1 || max_blen_y > PEL_MAXB) return hipErrorInvalidValue;
    FsLds ly;
    const size_t lds = fast_subpel_lds(max_blen_x, max_blen_y, ly);
    if (lds > kFastSubpelLdsMax) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_fast_subpel, dim3((total + 3) / 4), dim3(256), lds, st, jobs, tasks, ntasks, total, ly);
    return hipGetLastError();
}

hipError_t launch_fast_l2(const JobDesc *jobs, const Task *tasks, int ntasks, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fast_l2, dim3((total + 255) / 256), dim3(256), 0, st, jobs, tasks, ntasks, total);
    return hipGetLastError();
}

hipError_t launch_fast_mode(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int max_x, int max_y,
                            hipStream_t st) {
    if (total <= 0) return hipSuccess;
    int pd;
    const size_t lds = fast_mode_lds(max_x, max_y, pd);
    if (max_x < 1 || max_y < 1 || lds > kFastModeLdsMax) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_fast_mode, dim3((total + 3) / 4), dim3(256), lds, st, jobs, tasks, ntasks, total, pd, max_y);
    return hipGetLastError();
}

hipError_t launch_fast_eval(const JobDesc *jobs, const Task *tasks, int ntasks, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fast_eval, dim3((total + 3) / 4), dim3(256), 0, st, jobs, tasks, ntasks, total);
    return hipGetLastError();
}

}  // namespace dme
