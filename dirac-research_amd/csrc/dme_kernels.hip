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

#include <climits>
#include <cstdint>

#include "dme_device.h"

namespace dme {

#define WAVE 64

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

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

__device__ __forceinline__ int wave_sum(int v) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
    return v;
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
// One thread per output sample; the 12 short row-buffer values it needs are
// recomputed from the column taps.  Index clamping to [0, 2*len-1] is the
// reference's edge handling for len >= 12 (host-checked).
// ---------------------------------------------------------------------------
__constant__ int c_down_taps[6] = {86, 46, 4, -8, -4, 4};

__global__ __launch_bounds__(256) void k_downconvert(const int16_t *__restrict__ in, int in_w,
                                                     int16_t *__restrict__ out, int ow, int oh) {
    const int cx = blockIdx.x * 64 + (threadIdx.x & 63);
    const int cy = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (cx >= ow || cy >= oh) return;
    const int xlen = 2 * ow, ylen = 2 * oh;
    const int y = 2 * cy, x = 2 * cx;
    int rows[12];
    for (int k = 0; k < 6; ++k) {
        rows[5 - k] = clampi(y - k, 0, ylen - 1);
        rows[6 + k] = clampi(y + 1 + k, 0, ylen - 1);
    }
    int rb[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        const int col = clampi(x - 5 + q, 0, xlen - 1);
        int sum = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k)
            sum += ((int)in[(size_t)rows[5 - k] * in_w + col] + (int)in[(size_t)rows[6 + k] * in_w + col]) *
                   c_down_taps[k];
        rb[q] = (int16_t)((sum + 128) >> 8);
    }
    int sum = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) sum += (rb[5 - k] + rb[6 + k]) * c_down_taps[k];
    out[(size_t)cy * ow + cx] = (int16_t)((sum + 128) >> 8);
}

// ---------------------------------------------------------------------------
// K2: UpConverter::DoUpConverter (upconvert.cpp:55-200), clip [-128,127].
// One thread per source sample writes its 2x2 quad of the up plane.  The
// accumulator is a short in the reference; |sum| <= 16 + 32*255 fits, so int
// arithmetic gives the same value before the (short) >> 5.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int up_tap8(const int s[8]) {
    const int sum = 16 + (s[3] + s[4]) * 21 + (s[2] + s[5]) * -7 + (s[1] + s[6]) * 3 + (s[0] + s[7]) * -1;
    return clampi(sum >> 5, -128, 127);
}

__global__ __launch_bounds__(256) void k_upconvert(const int16_t *__restrict__ in, int W, int H,
                                                   int16_t *__restrict__ up) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    int rowi[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) rowi[t] = clampi(y - 3 + t, 0, H - 1);
    int ev[8], vt[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int xx = clampi(x - 3 + t, 0, W - 1);
        int s[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) s[q] = in[(size_t)rowi[q] * W + xx];
        ev[t] = s[3];
        vt[t] = up_tap8(s);
    }
    const size_t UW = 2 * (size_t)W;
    up[(2 * y) * UW + 2 * x] = (int16_t)ev[3];
    up[(2 * y) * UW + 2 * x + 1] = (int16_t)up_tap8(ev);
    up[(2 * y + 1) * UW + 2 * x] = (int16_t)vt[3];
    up[(2 * y + 1) * UW + 2 * x + 1] = (int16_t)up_tap8(vt);
}

// ---------------------------------------------------------------------------
// Pel search, bulk part: per block, the first minimum over the zero window
// and the guide window (the lists that do not depend on the prediction;
// pixel_match.cpp:304, 351-364).  One wave per block, four per workgroup.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pel_bulk(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                  int ntasks, int total, int level) {
    const int gi = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (gi >= total) return;
    const Task tk = tasks[find_task(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    const int r = tk.ref, local = gi - tk.offset;
    const int X = J.lxb[level];
    const int xb = local % X, yb = local / X;
    const int W = J.lw[level], H = J.lh[level];
    const Blk b = block_limits(J.blen[2], J.bsep[2], J.boff[2], W, H, xb, yb);
    int best_sad = 10000000;
    dme_mv best_mv;
    best_mv.x = 0;
    best_mv.y = 0;
    if (!b.empty()) {
        const int rad = min(level + 1, 5);
        const int n = (2 * rad + 1) * (2 * rad + 1);
        dme_mv g;
        g.x = 0;
        g.y = 0;
        int nc = n;
        if (level < J.depth) {
            const int GX = J.lxb[level + 1], GY = J.lyb[level + 1];
            const dme_mv gv = J.lvec[r][level + 1][(size_t)clampi(yb >> 1, 0, GY - 1) * GX + clampi(xb >> 1, 0, GX - 1)];
            g.x = gv.x * 2;
            g.y = gv.y * 2;
            nc = 2 * n;
        }
        const unsigned long long key = wave_pel_argmin(J.pic[level], J.ref[r][level], W, H, b, nc, [&](int c) {
            dme_mv m;
            if (c < n) {
                m = win_off(c, rad);
            } else {
                m = win_off(c - n, rad);
                m.x += g.x;
                m.y += g.y;
            }
            return m;
        });
        const int c = (int)(key & 0xffffffffu);
        best_sad = (int)(key >> 32);
        if (c < n) {
            best_mv = win_off(c, rad);
        } else {
            best_mv = win_off(c - n, rad);
            best_mv.x += g.x;
            best_mv.y += g.y;
        }
    }
    if ((threadIdx.x & 63) == 0) {
        J.bsad[r][local] = best_sad;
        J.bmv[r][local] = best_mv;
    }
}

// ---------------------------------------------------------------------------
// Pel search, wavefront part (PixelMatcher::MatchPic/DoBlock,
// pixel_match.cpp:242-380; BlockMatcher::FindBestMatchPel,
// block_match.cpp:242-275).  One workgroup per (picture, reference).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_pel_wave(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                   int level) {
    const Task tk = tasks[blockIdx.x];
    const JobDesc &J = jobs[tk.job];
    if (level > J.depth || tk.ref >= J.nrefs) return;  // uniform per workgroup
    const int r = tk.ref;
    const int X = J.lxb[level], Y = J.lyb[level];
    const int W = J.lw[level], H = J.lh[level];
    const int rad = min(level + 1, 5);
    const int n = (2 * rad + 1) * (2 * rad + 1);
    dme_mv *__restrict__ V = J.lvec[r][level];
    dme_cost *__restrict__ C = J.lpc[r][level];
    const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const int lane = threadIdx.x & 63;
    const int T = (X - 1) + 2 * (Y - 1);
    for (int t = 0; t <= T; ++t) {
        const int ylo = t - (X - 1) > 0 ? (t - (X - 1) + 1) >> 1 : 0;
        const int yhi = min(Y - 1, t >> 1);
        for (int yb = ylo + wave; yb <= yhi; yb += nwaves) {
            const int xb = t - 2 * yb;
            const size_t k = (size_t)yb * X + xb;
            // prediction rules (pixel_match.cpp:314-347)
            dme_mv pred;
            if (yb == 0) {
                if (xb == 0) { pred.x = 0; pred.y = 0; }
                else pred = V[xb - 1];
            } else if (xb == 0) {
                pred = V[k - X];
            } else if (xb < X - 1) {
                pred = med3(V[k - 1], V[k - X], V[k - X + 1]);
            } else {
                pred = mean2(V[k - X], V[k - 1]);
            }
            const Blk b = block_limits(J.blen[2], J.bsep[2], J.boff[2], W, H, xb, yb);
            int best_sad = J.bsad[r][k];
            dme_mv best = J.bmv[r][k];
            if (!b.empty()) {
                const unsigned long long key =
                    wave_pel_argmin(J.pic[level], J.ref[r][level], W, H, b, n, [&](int c) {
                        dme_mv m = win_off(c, rad);
                        m.x += pred.x;
                        m.y += pred.y;
                        return m;
                    });
                const int s = (int)(key >> 32);
                if (s < best_sad) {
                    best_sad = s;
                    best = win_off((int)(key & 0xffffffffu), rad);
                    best.x += pred.x;
                    best.y += pred.y;
                }
            }
            if (lane == 0) {
                V[k] = best;
                dme_cost c;
                c.SAD = (float)best_sad;
                c.mvcost = (float)(int16_t)norm1(best.x - pred.x, best.y - pred.y);
                c.total = c.SAD + 0.0f * c.mvcost;  // SetTotal(0.0)
                C[k] = c;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Up-sampled SADs (BlockDiffHalfPel / QuarterPel, me_utils.cpp:410-867;
// MMX simple_block_diff_up_mmx_4, me_utils_mmx.cpp:201-533).
// step 1: mv in half pel, sample up[2y+mv.y][2x+mv.x]; step 2: quarter pel,
// rounded average of 1, 2 or 4 half-pel samples.  Clamped reads equal both
// reference paths' values.
// ---------------------------------------------------------------------------
struct UpStart {
    int rsx, rsy, rx, ry;
    bool bounds;
};

__device__ __forceinline__ UpStart up_start(const Blk &b, dme_mv mv, int step, int UW, int UH) {
    UpStart s;
    if (step == 1) {
        s.rsx = (b.xp << 1) + mv.x; s.rsy = (b.yp << 1) + mv.y; s.rx = 0; s.ry = 0;
    } else {
        s.rsx = (b.xp << 1) + (mv.x >> 1); s.rsy = (b.yp << 1) + (mv.y >> 1);
        s.rx = mv.x & 1; s.ry = mv.y & 1;
    }
    s.bounds = s.rsx < 0 || s.rsx + (b.xl << 1) >= UW || s.rsy < 0 || s.rsy + (b.yl << 1) >= UH;
    return s;
}

__device__ __forceinline__ int up_val(const int16_t *__restrict__ up, int UW, int UH, const UpStart &s, int dy, int dx) {
    const int uy = s.rsy + 2 * dy, ux = s.rsx + 2 * dx;
    const int by = clampi(uy, 0, UH - 1), bx = clampi(ux, 0, UW - 1);
    const int a = up[(size_t)by * UW + bx];
    if (!(s.rx | s.ry)) return a;
    const int by1 = clampi(uy + 1, 0, UH - 1), bx1 = clampi(ux + 1, 0, UW - 1);
    const int w0 = (2 - s.rx) * (2 - s.ry), w1 = s.rx * (2 - s.ry), w2 = (2 - s.rx) * s.ry, w3 = s.rx * s.ry;
    return (w0 * a + w1 * up[(size_t)by * UW + bx1] + w2 * up[(size_t)by1 * UW + bx] + w3 * up[(size_t)by1 * UW + bx1] +
            2) >> 2;
}

// Integer SADs of up to 4 candidates against the up plane; wave-uniform.
__device__ void wave_up_sads(const int16_t *__restrict__ pic, int W, const int16_t *__restrict__ up, int UW, int UH,
                             const Blk &b, int step, const dme_mv *cands, int nc, int *out) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int R = pick_group(nc, b.yl);
    const int per = WAVE / R, g = lane / R, sub = lane - g * R;
    for (int base = 0; base < nc; base += per) {
        const int c = base + g;
        int s = 0;
        if (c < nc) {
            const UpStart st = up_start(b, cands[c], step, UW, UH);
            for (int row = sub; row < b.yl; row += R) {
                const int16_t *pr = pic + (size_t)(b.yp + row) * W + b.xp;
                for (int x = 0; x < b.xl; ++x) s += abs(up_val(up, UW, UH, st, row, x) - (int)pr[x]);
            }
        }
        for (int m = 1; m < R; m <<= 1) s += __shfl_xor(s, m, WAVE);
        for (int q = 0; q < per && base + q < nc; ++q) {
            const int v = __shfl(s, q * R, WAVE);
            out[base + q] = v;
        }
    }
}

// Per-pixel float accumulation of the bounds-checked path (and of the plain
// C build), raster order (me_utils.cpp:532-541, 843-867).  Every lane runs it
// on the same addresses, so the loads are broadcast.
__device__ float chain_total(float start, const int16_t *__restrict__ pic, int W, const int16_t *__restrict__ up,
                             int UW, int UH, const Blk &b, const UpStart &st) {
    float sum = start;
    for (int dy = 0; dy < b.yl; ++dy) {
        const int16_t *pr = pic + (size_t)(b.yp + dy) * W + b.xp;
        for (int dx = 0; dx < b.xl; ++dx) sum += (float)abs(up_val(up, UW, UH, st, dy, dx) - (int)pr[dx]);
    }
    return sum;
}

// The bail-out Diff with a start cost: update best when the candidate's
// total is strictly lower.  MMX semantics add the integer SAD to the start
// once (me_utils_mmx.cpp:270-278).
__device__ __forceinline__ void try_candidate(const JobDesc &J, const int16_t *up, int UW, int UH, const Blk &b,
                                              dme_mv mv, int step, int isad, float mvcost, float lambda,
                                              dme_cost &best, dme_mv &best_mv) {
    const float start = mvcost * lambda;
    const UpStart st = up_start(b, mv, step, UW, UH);
    float total;
    if (st.bounds || J.fp_sem == DME_FP_C)
        total = chain_total(start, J.pic[0], J.W, up, UW, UH, b, st);
    else
        total = (float)isad + start;
    if (total < best.total) {
        best_mv = mv;
        best.total = total;
        best.mvcost = mvcost;
        best.SAD = total - start;
    }
}

// GetVarUp (block_match.cpp:233-240) through its short return type
__device__ __forceinline__ float var_up(dme_mv pred, dme_mv m) {
    const int a = norm1(m.x - pred.x, m.y - pred.y), c = norm1(m.x, m.y);
    return (float)(int16_t)(a < c ? a : c);
}

// ---------------------------------------------------------------------------
// Sub-pel refinement (SubpelRefine::MatchPic/DoBlock/GetPred,
// me_subpel.cpp:101-176; BlockMatcher::RefineMatchSubp,
// block_match.cpp:327-467).  One workgroup per (picture, reference),
// wavefront t = x + 2y.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_subpel_wave(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks) {
    const Task tk = tasks[blockIdx.x];
    const JobDesc &J = jobs[tk.job];
    if (tk.ref >= J.nrefs || J.prec == 0) return;
    const int r = tk.ref, prec = J.prec;
    const int X = J.xnb, Y = J.ynb, W = J.W, H = J.H;
    const int UW = 2 * W, UH = 2 * H;
    const int16_t *up = J.up[r];
    dme_mv *__restrict__ V = J.vec[r];
    dme_cost *__restrict__ C = J.pc[r];
    const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6, lane = threadIdx.x & 63;
    const int T = (X - 1) + 2 * (Y - 1);
    for (int t = 0; t <= T; ++t) {
        const int ylo = t - (X - 1) > 0 ? (t - (X - 1) + 1) >> 1 : 0;
        const int yhi = min(Y - 1, t >> 1);
        for (int yb = ylo + wave; yb <= yhi; yb += nwaves) {
            const int xb = t - 2 * yb;
            const size_t k = (size_t)yb * X + xb;
            // GetPred: median of the available left, top-left, top, top-right
            dme_mv nb[4];
            int nn = 0;
            if (xb > 0) nb[nn++] = V[k - 1];
            if (xb > 0 && yb > 0) nb[nn++] = V[k - X - 1];
            if (yb > 0) nb[nn++] = V[k - X];
            if (yb > 0 && xb + 1 < X) nb[nn++] = V[k - X + 1];
            dme_mv pred;
            if (nn == 0) { pred.x = 0; pred.y = 0; }
            else if (nn == 1) pred = nb[0];
            else if (nn == 2) pred = mean2(nb[0], nb[1]);
            else if (nn == 3) pred = med3(nb[0], nb[1], nb[2]);
            else {
                int sx = 0, sy = 0, mxx = nb[0].x, mxy = nb[0].y, mnx = nb[0].x, mny = nb[0].y;
                for (int i = 0; i < 4; ++i) {
                    sx += nb[i].x; sy += nb[i].y;
                    mxx = max(mxx, nb[i].x); mxy = max(mxy, nb[i].y);
                    mnx = min(mnx, nb[i].x); mny = min(mny, nb[i].y);
                }
                pred.x = (sx - (mxx + mnx) + 1) >> 1;
                pred.y = (sy - (mxy + mny) + 1) >> 1;
            }
            const float lambda = J.lam[k];
            const Blk b = block_limits(J.blen[2], J.bsep[2], J.boff[2], W, H, xb, yb);
            const dme_mv pel = V[k];
            dme_cost cost = C[k];
            dme_mv pel_up;
            pel_up.x = pel.x << prec;
            pel_up.y = pel.y << prec;
            cost.mvcost = var_up(pred, pel_up);
            cost.total = cost.SAD + lambda * cost.mvcost;
            const float lim = (float)(2 * b.xl * b.yl);
            dme_mv out_mv;
            dme_cost out_cost;
            if (cost.SAD < lim) {
                out_mv = pel_up;
                out_cost = cost;
            } else {
                dme_cost pcost;
                pcost.mvcost = 0.0f;
                int ps = 0;
                if (!b.empty()) wave_up_sads(J.pic[0], W, up, UW, UH, b, prec, &pred, 1, &ps);
                pcost.SAD = (float)ps;
                pcost.total = pcost.SAD;
                if (pcost.SAD < lim) {
                    out_mv = pred;
                    out_cost = pcost;
                } else {
                    dme_cost best = cost;
                    dme_mv bmv = pel;
                    bool bailed = false;
                    for (int i = 1; i <= prec; ++i) {
                        bmv.x <<= 1;
                        bmv.y <<= 1;
                        dme_mv tb = bmv;
                        const int sh = prec - i;
                        if (!b.empty()) {
                            dme_mv cd[4];
                            cd[0].x = bmv.x - 1; cd[0].y = bmv.y;
                            cd[1].x = bmv.x + 1; cd[1].y = bmv.y;
                            cd[2].x = bmv.x; cd[2].y = bmv.y - 1;
                            cd[3].x = bmv.x; cd[3].y = bmv.y + 1;
                            int sd[4];
                            wave_up_sads(J.pic[0], W, up, UW, UH, b, i, cd, 4, sd);
                            for (int q = 0; q < 4; ++q) {
                                dme_mv cs;
                                cs.x = cd[q].x << sh;
                                cs.y = cd[q].y << sh;
                                try_candidate(J, up, UW, UH, b, cd[q], i, sd[q], var_up(pred, cs), lambda, best, tb);
                            }
                            int n2 = 0;
                            if (tb.x != bmv.x) {
                                cd[0].x = tb.x; cd[0].y = tb.y - 1;
                                cd[1].x = tb.x; cd[1].y = tb.y + 1;
                                n2 = 2;
                            } else if (tb.y != bmv.y) {
                                cd[0].x = tb.x - 1; cd[0].y = tb.y;
                                cd[1].x = tb.x + 1; cd[1].y = tb.y;
                                n2 = 2;
                            }
                            if (n2) {
                                wave_up_sads(J.pic[0], W, up, UW, UH, b, i, cd, 2, sd);
                                for (int q = 0; q < 2; ++q) {
                                    dme_mv cs;
                                    cs.x = cd[q].x << sh;
                                    cs.y = cd[q].y << sh;
                                    try_candidate(J, up, UW, UH, b, cd[q], i, sd[q], var_up(pred, cs), lambda, best, tb);
                                }
                            }
                        }
                        bmv = tb;
                        if ((double)best.total > 1.1 * (double)pcost.total) {
                            bailed = true;
                            break;
                        }
                    }
                    if (bailed) {
                        out_mv = pred;
                        out_cost = pcost;
                    } else {
                        out_mv = bmv;
                        out_cost = best;
                    }
                }
            }
            if (lane == 0) {
                V[k] = out_mv;
                C[k] = out_cost;
            }
        }
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

// BiBlock{Half,Quarter}Pel::Diff (me_utils.cpp:1172-1543): sum of
// |((2 pic - r1) - r2) >> 1| with short intermediates.
__device__ int wave_bi_sad(const JobDesc &J, const Blk &b, dme_mv m1, dme_mv m2) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int UW = 2 * J.W, UH = 2 * J.H;
    const UpStart s1 = up_start(b, m1, J.step, UW, UH);
    const UpStart s2 = up_start(b, m2, J.step, UW, UH);
    int s = 0;
    const int n = b.xl * b.yl;
    for (int i = lane; i < n; i += WAVE) {
        const int dy = i / b.xl, dx = i - dy * b.xl;
        const int p = J.pic[0][(size_t)(b.yp + dy) * J.W + b.xp + dx];
        const int16_t diff = (int16_t)((p << 1) - up_val(J.up[0], UW, UH, s1, dy, dx));
        const int16_t temp = (int16_t)(((int)diff - up_val(J.up[1], UW, UH, s2, dy, dx)) >> 1);
        s += abs((int)temp);
    }
    return wave_sum(s);
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
// parallel: it only reads the block's own refined vectors and costs.
__global__ __launch_bounds__(256) void k_mode_l2(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                 int ntasks, int total) {
    const int gi = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (gi >= total) return;
    const Task tk = tasks[find_task(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
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
        bc.SAD = b.empty() ? 0.0f : (float)wave_bi_sad(J, b, J.vec[0][k], J.vec[1][k]);
        bc.total = bc.SAD + loc_lambda * bc.mvcost;
        bc.total *= lf;
        uc = bc.total + mode_cost;
        if (uc < min_cost) {
            mode = DME_REF1AND2;
            min_cost = uc;
            best_sad = bc.SAD;
        }
        if (lane == 0) J.bi[k] = bc;
    }
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
        if (lane == 0) {
            J.intra[k] = ic;
            J.dc[0][k] = (int16_t)dc;
        }
    }
    if (lane == 0) {
        J.mode[k] = mode;
        J.ucost[k] = min_cost;
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

// ModeDecider::DoME for one unit and reference (me_mode_decn.cpp:279-364,
// BlockMatcher::FindBestMatchSubp, block_match.cpp:277-325).  cands are the
// four guide vectors; duplicates are dropped as AddNewVlist does.
__device__ void unit_me(const JobDesc &J, int level, int xu, int yu, int r, const dme_mv g[4], float lambda,
                        dme_mv &out_mv, dme_cost &out_cost) {
    dme_mv cd[4];
    int nc = 0;
    for (int q = 0; q < 4; ++q) {
        bool dup = false;
        for (int s = 0; s < nc; ++s) dup = dup || (cd[s].x == g[q].x && cd[s].y == g[q].y);
        if (!dup) cd[nc++] = g[q];
    }
    const dme_mv pred = mode_pred(J.vec[r], J.xnb, xu << (2 - level), yu << (2 - level));
    const Blk b = block_limits(J.blen[level], J.bsep[level], J.boff[level], J.W, J.H, xu, yu);
    dme_cost best;
    best.SAD = 0.0f;
    best.mvcost = 0.0f;
    best.total = 100000000.0f;
    dme_mv bmv;
    bmv.x = 0;
    bmv.y = 0;
    if (!b.empty()) {
        const int UW = 2 * J.W, UH = 2 * J.H;
        int sd[4];
        wave_up_sads(J.pic[0], J.W, J.up[r], UW, UH, b, J.step, cd, nc, sd);
        for (int q = 0; q < nc; ++q)
            try_candidate(J, J.up[r], UW, UH, b, cd[q], J.step, sd[q], var_up(pred, cd[q]), lambda, best, bmv);
    }
    out_mv = bmv;
    out_cost = best;
}

// DoUnitDecn for levels 1 and 0 (me_mode_decn.cpp:368-441; no intra).
__device__ float unit_decn(const JobDesc &J, int level, int xu, int yu, float loc_lambda, dme_mv v[2], dme_cost c[2],
                           int &mode) {
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
        const Blk b = block_limits(J.blen[level], J.bsep[level], J.boff[level], J.W, J.H, xu, yu);
        dme_cost bc;
        bc.mvcost = c[0].mvcost + c[1].mvcost;
        bc.SAD = b.empty() ? 0.0f : (float)wave_bi_sad(J, b, v[0], v[1]);
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
// superblock cost after level 1 never exceeds the level-2 cost.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_mode_wave(const JobDesc *__restrict__ jobs) {
    const JobDesc &J = jobs[blockIdx.x];
    const int XS = J.xsb, YS = J.ysb, X2 = J.xnb;
    const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6, lane = threadIdx.x & 63;
    const int nr = J.nrefs;
    const int X1 = J.xnb / 2;
    for (int t = 0; t <= (XS - 1) + (YS - 1); ++t) {
        const int ylo = t - (XS - 1) > 0 ? t - (XS - 1) : 0;
        const int yhi = min(YS - 1, t);
        for (int ys = ylo + wave; ys <= yhi; ys += nwaves) {
            const int xs = t - ys;
            const size_t s = (size_t)ys * XS + xs;
            // level 2: sum of the block costs in raster order
            float sbc = 0.0f;
            for (int j = 0; j < 4; ++j)
                for (int i = 0; i < 4; ++i) sbc += J.ucost[(size_t)(4 * ys + j) * X2 + 4 * xs + i];
            int split = 2;
            // level 1
            dme_mv v1[4][2];
            int m1[4];
            float sb1 = 0.0f;
            for (int u = 0; u < 4; ++u) {
                const int xu = 2 * xs + (u & 1), yu = 2 * ys + (u >> 1);
                const float lam = J.lam1[(size_t)yu * X1 + xu];
                dme_cost c[2];
                for (int r = 0; r < nr; ++r) {
                    dme_mv g[4];
                    for (int q = 0; q < 4; ++q)
                        g[q] = J.vec[r][(size_t)(2 * yu + (q >> 1)) * X2 + 2 * xu + (q & 1)];
                    unit_me(J, 1, xu, yu, r, g, lam, v1[u][r], c[r]);
                }
                sb1 += unit_decn(J, 1, xu, yu, lam, v1[u], c, m1[u]);
            }
            if (sb1 <= sbc) {
                sbc = sb1;
                split = 1;
                if (lane < 16) {
                    const int u = ((lane >> 3) << 1) | ((lane >> 1) & 1);
                    const size_t d = (size_t)(4 * ys + (lane >> 2)) * X2 + 4 * xs + (lane & 3);
                    J.mode[d] = m1[u];
                    J.dc[0][d] = 0;
                    J.vec[0][d] = v1[u][0];
                    if (nr > 1) J.vec[1][d] = v1[u][1];
                }
            }
            // level 0
            {
                const float lam = J.lam0[s];
                dme_mv v0[2];
                dme_cost c[2];
                for (int r = 0; r < nr; ++r) {
                    dme_mv g[4];
                    for (int q = 0; q < 4; ++q) g[q] = v1[q][r];
                    unit_me(J, 0, xs, ys, r, g, lam, v0[r], c[r]);
                }
                int m0;
                const float sb0 = unit_decn(J, 0, xs, ys, lam, v0, c, m0);
                if (sb0 <= sbc) {
                    sbc = sb0;
                    split = 0;
                    if (lane < 16) {
                        const size_t d = (size_t)(4 * ys + (lane >> 2)) * X2 + 4 * xs + (lane & 3);
                        J.mode[d] = m0;
                        J.dc[0][d] = 0;
                        J.vec[0][d] = v0[0];
                        if (nr > 1) J.vec[1][d] = v0[1];
                    }
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
// K7: ModeDecider::SetDC (me_mode_decn.cpp:523-573): block DCs of Y, U, V,
// then the 3x3 [1 3 1; 3 0 3; 1 3 1]/16 smoothing; row 0 and the first and
// last columns are copied and the last row is left as it was.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_block_dc(const JobDesc *__restrict__ jobs, const Task *__restrict__ tasks,
                                                  int ntasks, int total) {
    const int gi = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (gi >= total) return;
    const Task tk = tasks[find_task(tasks, ntasks, gi)];
    const JobDesc &J = jobs[tk.job];
    const int comp = tk.ref, k = gi - tk.offset;
    const int xb = k % J.xnb, yb = k / J.xnb;
    int dc = 0;
    if (comp == 0) {
        const Blk b = block_limits(J.blen[2], J.bsep[2], J.boff[2], J.W, J.H, xb, yb);
        if (!b.empty()) dc = wave_block_dc(J.pic[0], J.W, b);
    } else {
        const int16_t *p = comp == 1 ? J.cu : J.cv;
        const Blk b = block_limits(J.cblen, J.cbsep, J.cboff, J.CW, J.CH, xb, yb);
        if (p && !b.empty()) dc = wave_block_dc(p, J.CW, b);
    }
    if ((threadIdx.x & 63) == 0) J.dct[comp][k] = (int16_t)dc;
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

hipError_t launch_downconvert(const int16_t *in, int in_w, int16_t *out, int ow, int oh, hipStream_t st) {
    dim3 grid((ow + 63) / 64, (oh + 3) / 4);
    hipLaunchKernelGGL(k_downconvert, grid, dim3(256), 0, st, in, in_w, out, ow, oh);
    return hipGetLastError();
}

hipError_t launch_upconvert(const int16_t *in, int W, int H, int16_t *up, hipStream_t st) {
    dim3 grid((W + 63) / 64, (H + 3) / 4);
    hipLaunchKernelGGL(k_upconvert, grid, dim3(256), 0, st, in, W, H, up);
    return hipGetLastError();
}

hipError_t launch_pel_bulk(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int level, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pel_bulk, dim3((total + 3) / 4), dim3(256), 0, st, jobs, tasks, ntasks, total, level);
    return hipGetLastError();
}

hipError_t launch_pel_wave(const JobDesc *jobs, const Task *tasks, int ntasks, int level, int threads, hipStream_t st) {
    hipLaunchKernelGGL(k_pel_wave, dim3(ntasks), dim3(threads), 0, st, jobs, tasks, level);
    return hipGetLastError();
}

hipError_t launch_subpel_wave(const JobDesc *jobs, const Task *tasks, int ntasks, int threads, hipStream_t st) {
    hipLaunchKernelGGL(k_subpel_wave, dim3(ntasks), dim3(threads), 0, st, jobs, tasks);
    return hipGetLastError();
}

hipError_t launch_shift_vectors(const JobDesc *jobs, const Task *tasks, int ntasks, int total, int left,
                                hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_shift_vectors, dim3((total + 255) / 256), dim3(256), 0, st, jobs, tasks, ntasks, total, left);
    return hipGetLastError();
}

hipError_t launch_mode_l2(const JobDesc *jobs, const Task *tasks, int ntasks, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mode_l2, dim3((total + 3) / 4), dim3(256), 0, st, jobs, tasks, ntasks, total);
    return hipGetLastError();
}

hipError_t launch_mode_wave(const JobDesc *jobs, int njobs, int threads, hipStream_t st) {
    hipLaunchKernelGGL(k_mode_wave, dim3(njobs), dim3(threads), 0, st, jobs);
    return hipGetLastError();
}

hipError_t launch_block_dc(const JobDesc *jobs, const Task *tasks, int ntasks, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_block_dc, dim3((total + 3) / 4), dim3(256), 0, st, jobs, tasks, ntasks, total);
    return hipGetLastError();
}

hipError_t launch_dc_smooth(const JobDesc *jobs, const Task *tasks, int ntasks, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dc_smooth, dim3((total + 255) / 256), dim3(256), 0, st, jobs, tasks, ntasks, total);
    return hipGetLastError();
}

}  // namespace dme
