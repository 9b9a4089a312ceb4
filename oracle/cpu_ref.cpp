// cpu_ref.cpp -- CPU oracle for the Dirac encoder motion-estimation path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain scalar C++, what
// the reference (MOXfiles/dirac-research, release 1.0.2) computes on its live
// ME path, so that the HIP implementation can be checked against it.  It is
// built into oracle/liboracle.so, which only tests/, smoke() and the
// cpu_baseline leg of bench.py load.  Nothing in the product links it.
//
// Each function names the reference lines it follows.  Float work is done in
// the same types and order as the reference (float / double / long double),
// compiled with -ffp-contract=off so no multiply-add is fused.
//
// Parity pins (see tests/): every stage is compared bit-exactly with the
// reference compiled from its own sources (oracle/_ref, via ref_harness.cpp)
// on seeded inputs, and with committed golden fixtures produced that way.
// K1/K2 are additionally pinned by closed forms (constant in -> constant out),
// the SAD kernels by brute force on tiny inputs and the search by
// translation-recovery invariants.

#include "cpu_ref.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string &m) {
    g_err = m;
    return DME_ERR_INVALID;
}

// ---------------------------------------------------------------------------
// Basic types
// ---------------------------------------------------------------------------

struct Plane {
    int w = 0, h = 0;
    std::vector<int16_t> d;
    Plane() {}
    Plane(int w_, int h_) : w(w_), h(h_), d((size_t)w_ * h_, 0) {}
    int16_t at(int y, int x) const { return d[(size_t)y * w + x]; }
    int16_t &at(int y, int x) { return d[(size_t)y * w + x]; }
};

struct MV {
    int x = 0, y = 0;
    MV() {}
    MV(int a, int b) : x(a), y(b) {}
    bool operator==(const MV &o) const { return x == o.x && y == o.y; }
};

struct Cost {
    float SAD = 0.f, mvcost = 0.f, total = 0.f;
};

// OLBParams (common.h:817-870, common.cpp:155-162): offsets follow the
// setters, (len - sep) / 2.
struct OLB {
    int xblen = 0, yblen = 0, xbsep = 0, ybsep = 0;
    int xoff() const { return (xblen - xbsep) / 2; }
    int yoff() const { return (yblen - ybsep) / 2; }
};

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// BChk (common.h:1574-1579): clamp an index into [0, max-1].
inline int bchk(int num, int max) { return num < 0 ? 0 : (num >= max ? max - 1 : num); }

// Norm1 (motion.h:435-437), value returned through a ValueType (short) as
// GetVar / GetVarUp do (block_match.cpp:224-240).
inline int norm1(const MV &m) { return std::abs(m.x) + std::abs(m.y); }

// ---------------------------------------------------------------------------
// Geometry: PicturePredParams::SetBlockSizes (common.cpp:207-320),
// SequenceCompressor::SetMotionParameters (seq_compress.cpp:107-133),
// PixelMatcher::DoSearch depth (pixel_match.cpp:94-95) and
// MakeMEDataHierarchy (pixel_match.cpp:185-220).
// ---------------------------------------------------------------------------

struct Geo {
    int W, H, CW, CH;
    OLB l[3], c[3];
    int xsb, ysb, xnb, ynb;
    int depth;
    int lw[DME_MAX_DEPTH + 1], lh[DME_MAX_DEPTH + 1];
    int lxb[DME_MAX_DEPTH + 1], lyb[DME_MAX_DEPTH + 1];
    float level_factor[3];
};

int make_geo(const dme_params *p, int W, int H, Geo &g) {
    if (!p) return fail("params is NULL");
    if (p->chroma_format != 0) return fail("only 4:2:0 is supported");
    if (p->xbsep <= 0 || p->ybsep <= 0 || p->xblen <= 0 || p->yblen <= 0)
        return fail("block lengths and separations must be positive");
    if (p->mv_precision < 0 || p->mv_precision > 3)
        return fail("mv_precision must be 0, 1, 2 or 3");
    if (W < 24 || H < 24) return fail("picture too small");
    g.W = W; g.H = H; g.CW = W >> 1; g.CH = H >> 1;

    // SetBlockSizes, level 2 rationalisation (common.cpp:226-276).
    OLB b;
    b.xblen = p->xblen; b.yblen = p->yblen; b.xbsep = p->xbsep; b.ybsep = p->ybsep;
    int rem = b.xbsep % 4;
    if (rem != 0 || b.xbsep == 0) { b.xbsep += 4 - rem; b.xblen = b.xbsep + 4; }
    rem = b.ybsep % 4;
    if (rem != 0 || b.ybsep == 0) { b.ybsep += 4 - rem; b.yblen = b.ybsep + 4; }
    if (b.xblen % 4 != 0) b.xblen = b.xbsep + 4;
    if (b.yblen % 4 != 0) b.yblen = b.ybsep + 4;
    if (b.xbsep > b.xblen) b.xblen = b.xbsep + 4;
    if (b.ybsep > b.yblen) b.yblen = b.ybsep + 4;
    if (2 * b.xbsep < b.xblen) b.xblen = b.xbsep + 4;
    if (2 * b.ybsep < b.yblen) b.yblen = b.ybsep + 4;
    g.l[2] = b;
    // chroma, 4:2:0 factors of 2 (common.cpp:278-282)
    g.c[2].xbsep = b.xbsep / 2; g.c[2].xblen = b.xblen / 2;
    g.c[2].ybsep = b.ybsep / 2; g.c[2].yblen = b.yblen / 2;
    // levels 1 and 0: the y length adds the *x* separation (common.cpp:294-313)
    for (int k = 1; k >= 0; --k) {
        const OLB &u = g.l[k + 1];
        g.l[k].xbsep = u.xbsep * 2; g.l[k].xblen = u.xblen + u.xbsep;
        g.l[k].ybsep = u.ybsep * 2; g.l[k].yblen = u.yblen + u.xbsep;
        const OLB &cu = g.c[k + 1];
        g.c[k].xbsep = cu.xbsep * 2; g.c[k].xblen = cu.xblen + cu.xbsep;
        g.c[k].ybsep = cu.ybsep * 2; g.c[k].yblen = cu.yblen + cu.xbsep;
    }
    // SetMotionParameters (seq_compress.cpp:124-132)
    g.xsb = (W + g.l[0].xbsep - 1) / g.l[0].xbsep;
    g.ysb = (H + g.l[0].ybsep - 1) / g.l[0].ybsep;
    g.xnb = 4 * g.xsb;
    g.ynb = 4 * g.ysb;

    // depth (pixel_match.cpp:94-95); full search has no hierarchy (:91, 140-155)
    if (p->full_search) {
        if (p->x_range_me < 0 || p->y_range_me < 0) return fail("full-search ranges must not be negative");
        g.depth = 0;
    } else if (p->depth < 0) {
        g.depth = (int)std::min(std::log(((double)W) / 12.0) / std::log(2.0),
                                std::log(((double)H) / 12.0) / std::log(2.0));
    } else {
        g.depth = p->depth;
    }
    if (!p->full_search && (g.depth < 1 || g.depth > DME_MAX_DEPTH)) return fail("pyramid depth out of range");

    // pyramid sizes from full resolution (pixel_match.cpp:166-172) and
    // MEData grids (pixel_match.cpp:185-220)
    g.lw[0] = W; g.lh[0] = H; g.lxb[0] = g.xnb; g.lyb[0] = g.ynb;
    int sf = 1, xn = g.xnb, yn = g.ynb;
    for (int i = 1; i <= g.depth; ++i) {
        sf *= 2;
        g.lw[i] = W / sf; g.lh[i] = H / sf;
        xn >>= 1; yn >>= 1;
        if (g.lw[i] % b.xbsep != 0) xn++;
        if (g.lh[i] % b.ybsep != 0) yn++;
        g.lxb[i] = xn; g.lyb[i] = yn;
        // DoDownConvert's three-part loops need 12 output rows/columns
        if (g.lw[i] < 12 || g.lh[i] < 12) return fail("pyramid level smaller than 12 samples");
        if (xn < 2) return fail("pyramid level has fewer than 2 block columns");
    }
    // ModeDecider level factors (me_mode_decn.cpp:79-85): float division
    g.level_factor[0] = float(16 * g.l[2].xblen * g.l[2].yblen) /
                        float(g.l[0].xblen * g.l[0].yblen);
    g.level_factor[1] = float(4 * g.l[2].xblen * g.l[2].yblen) /
                        float(g.l[1].xblen * g.l[1].yblen);
    g.level_factor[2] = 1.0f;
    return DME_OK;
}

// BlockDiffParams::SetBlockLimits (me_utils.cpp:58-78).  Lengths may come out
// zero or negative for blocks off the picture; every Diff then returns early.
struct Blk {
    int xp, yp, xl, yl;
    int xend() const { return xp + xl; }
    int yend() const { return yp + yl; }
    bool empty() const { return xl <= 0 || yl <= 0; }
};

Blk block_limits(const OLB &bp, int W, int H, int xb, int yb) {
    Blk r;
    const int lx = xb * bp.xbsep - bp.xoff();
    const int ly = yb * bp.ybsep - bp.yoff();
    r.xp = std::max(lx, 0);
    r.yp = std::max(ly, 0);
    r.xl = bp.xblen - r.xp + lx;
    r.yl = bp.yblen - r.yp + ly;
    if (r.xp + r.xl - 1 > W - 1) r.xl = W - r.xp;
    if (r.yp + r.yl - 1 > H - 1) r.yl = H - r.yp;
    return r;
}

// ---------------------------------------------------------------------------
// K1: DownConverter::DoDownConvert + RowLoop (downconvert.cpp:51-204; the
// MMX vertical pass downconvert_mmx.cpp:61-245 computes the same sums with
// pmaddwd).  12-tap symmetric half-band filter (downconvert.h:81-89),
// vertical into a short row buffer, then horizontal, +128 >> 8 each.  The
// three-part loops clamp indices into [0, 2*out_len - 1]; with out_len >= 12
// (checked) that is exactly a clamp-to-edge of every tap index.
// ---------------------------------------------------------------------------

const int kDownTaps[6] = {86, 46, 4, -8, -4, 4};

void downconvert(const Plane &in, Plane &out) {
    const int xlen = 2 * out.w, ylen = 2 * out.h;
    std::vector<int16_t> rowbuf(xlen);
    for (int cy = 0; cy < out.h; ++cy) {
        const int y = 2 * cy;
        for (int x = 0; x < xlen; ++x) {
            int sum = 0;
            for (int k = 0; k < 6; ++k)
                sum += (in.at(clampi(y - k, 0, ylen - 1), x) +
                        in.at(clampi(y + 1 + k, 0, ylen - 1), x)) * kDownTaps[k];
            sum += 128;
            rowbuf[x] = (int16_t)(sum >> 8);
        }
        for (int cx = 0; cx < out.w; ++cx) {
            const int x = 2 * cx;
            int sum = 0;
            for (int k = 0; k < 6; ++k)
                sum += (rowbuf[clampi(x - k, 0, xlen - 1)] +
                        rowbuf[clampi(x + 1 + k, 0, xlen - 1)]) * kDownTaps[k];
            sum += 128;
            out.at(cy, cx) = (int16_t)(sum >> 8);
        }
    }
}

// ---------------------------------------------------------------------------
// K2: UpConverter::DoUpConverter + RowLoop (upconvert.cpp:55-200) with
// min/max = -128/127 (enc_picture.cpp:148-150).  Even samples copied; odd
// ones from the 8-tap (21,-7,3,-1) filter, accumulated in a short, +16 >> 5,
// clipped.  Rows first (from the source), then columns of every up row.
// The leading/middle/trailing loops are exactly clamp-to-edge of the source
// index for w, h >= 8 (checked by callers).
// ---------------------------------------------------------------------------

const int kUpTaps[4] = {21, -7, 3, -1};

inline int16_t up_filter(const int16_t *s) {  // s[0..7]: pairs (3-t, 4+t)
    int16_t sum = 16;
    for (int t = 0; t < 4; ++t) sum = (int16_t)(sum + (s[3 - t] + s[4 + t]) * kUpTaps[t]);
    sum = (int16_t)(sum >> 5);
    return (int16_t)clampi(sum, -128, 127);
}

void upconvert(const Plane &in, Plane &up) {
    const int W = in.w, H = in.h;
    up = Plane(2 * W, 2 * H);
    int16_t s[8];
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            up.at(2 * y, 2 * x) = in.at(y, x);
            for (int t = 0; t < 8; ++t) s[t] = in.at(clampi(y - 3 + t, 0, H - 1), x);
            up.at(2 * y + 1, 2 * x) = up_filter(s);
        }
    }
    for (int r = 0; r < 2 * H; ++r) {
        for (int x = 0; x < W; ++x) {
            for (int t = 0; t < 8; ++t) s[t] = up.at(r, 2 * clampi(x - 3 + t, 0, W - 1));
            up.at(r, 2 * x + 1) = up_filter(s);
        }
    }
}

// ---------------------------------------------------------------------------
// Candidate windows: AddNewVlist / AddVect (block_match.cpp:89-126, 170-193).
// Enumeration order: centre; (+i,0),(-i,0) for i=1..xr; then for j=1..yr,
// i=-xr..xr: (i,+j),(i,-j).  A vector already present in any earlier list
// is skipped and an empty list is dropped; neither changes the result of a
// strict-< first-wins argmin, so we keep the skip only for the SAD count.
// ---------------------------------------------------------------------------

void window(const MV &c, int xr, int yr, std::vector<MV> &out) {
    out.clear();
    out.push_back(c);
    for (int i = 1; i <= xr; ++i) {
        out.push_back(MV(c.x + i, c.y));
        out.push_back(MV(c.x - i, c.y));
    }
    for (int j = 1; j <= yr; ++j)
        for (int i = -xr; i <= xr; ++i) {
            out.push_back(MV(c.x + i, c.y + j));
            out.push_back(MV(c.x + i, c.y - j));
        }
}

// ---------------------------------------------------------------------------
// K3: PelBlockDiff (me_utils.cpp:138-282; MMX simple_block_diff_mmx_4 and
// bchk_simple_block_diff_mmx_4, me_utils_mmx.cpp:53-108, 1050-1132).  The
// integer SAD with clamp-to-edge ref addressing; the in-bounds and bounds-
// checked paths give the same integers and the row bail-outs never change
// which candidate wins.
// ---------------------------------------------------------------------------

int pel_sad(const Plane &pic, const Plane &ref, const Blk &b, const MV &mv) {
    int sum = 0;
    for (int j = b.yp; j < b.yend(); ++j)
        for (int i = b.xp; i < b.xend(); ++i)
            sum += std::abs(pic.at(j, i) - ref.at(bchk(j + mv.y, ref.h), bchk(i + mv.x, ref.w)));
    return sum;
}

// ---------------------------------------------------------------------------
// K4: up-sampled block SADs (me_utils.cpp:410-1170, MMX
// simple_block_diff_up_mmx_4, me_utils_mmx.cpp:201-533).
// prec_step 1 = BlockDiffHalfPel (mv in half pel), 2 = BlockDiffQuarterPel,
// 3 = BlockDiffEighthPel.  Sample value at quarter pel uses the rounded
// average of 1/2/4 half-pel samples; the bounds path's weighted form
// (2-rx)(2-ry)... +2 >> 2 equals it.  At eighth pel the start is mv >> 2 in
// the up plane and the weights (4-rx)(4-ry), rx(4-ry), (4-rx)ry, rx ry of
// the remainders rx, ry in 0..3, +8 >> 4 (me_utils.cpp:890-1011).
// ---------------------------------------------------------------------------

struct UpSample {
    int rsx, rsy;   // ref start in the up plane
    int rx, ry;     // quarter (0..1) or eighth (0..3) remainders (0 for half pel)
    bool eighth;    // eighth-pel weights
    bool bounds;    // true when the reference takes the bounds-checked path
};

UpSample up_start(const Blk &b, const MV &mv, int step, int upw, int uph) {
    UpSample s;
    s.eighth = step == 3;
    if (step == 1) {
        s.rsx = (b.xp << 1) + mv.x; s.rsy = (b.yp << 1) + mv.y; s.rx = 0; s.ry = 0;
    } else if (step == 2) {
        s.rsx = (b.xp << 1) + (mv.x >> 1); s.rsy = (b.yp << 1) + (mv.y >> 1);
        s.rx = mv.x & 1; s.ry = mv.y & 1;
    } else {
        s.rsx = (b.xp << 1) + (mv.x >> 2); s.rsy = (b.yp << 1) + (mv.y >> 2);
        s.rx = mv.x & 3; s.ry = mv.y & 3;
    }
    const int stopx = s.rsx + (b.xl << 1), stopy = s.rsy + (b.yl << 1);
    s.bounds = s.rsx < 0 || stopx >= upw || s.rsy < 0 || stopy >= uph;
    return s;
}

inline int up_value(const Plane &up, const UpSample &s, int dy, int dx) {
    const int uy = s.rsy + 2 * dy, ux = s.rsx + 2 * dx;
    const int by = bchk(uy, up.h), bx = bchk(ux, up.w);
    const int by1 = bchk(uy + 1, up.h), bx1 = bchk(ux + 1, up.w);
    const int one = s.eighth ? 4 : 2;
    const int w0 = (one - s.rx) * (one - s.ry), w1 = s.rx * (one - s.ry);
    const int w2 = (one - s.rx) * s.ry, w3 = s.rx * s.ry;
    const int v = w0 * up.at(by, bx) + w1 * up.at(by, bx1) + w2 * up.at(by1, bx) + w3 * up.at(by1, bx1);
    return s.eighth ? (v + 8) >> 4 : (v + 2) >> 2;
}

// Bail-out Diff with a start cost (BlockDiffHalfPel/QuarterPel::Diff with
// mvcost, me_utils.cpp:481-571, 686-867).  Returns true and fills best when
// the candidate's total beats best.total.  MMX semantics: in-bounds total =
// float(int SAD) + start, one rounding (me_utils_mmx.cpp:270-278); the
// bounds path (and the plain C build everywhere) adds per pixel in float.
// fp_sem FP_ONE_ROUNDING: float(int SAD) + start for every block (the fast
// form's own searches, which no reference path constrains).  Eighth pel has
// no MMX form: it adds per pixel in float in both builds (me_utils.cpp:
// 1013-1170).
constexpr int FP_ONE_ROUNDING = 2;

bool up_diff_update(const Plane &pic, const Plane &up, const Blk &b, const MV &mv, int step,
                    float mvcost, float lambda, int fp_sem, Cost &best, MV &best_mv,
                    double &terms) {
    if (b.empty()) return false;
    const UpSample s = up_start(b, mv, step, up.w, up.h);
    const float start_val = mvcost * lambda;
    terms += (double)b.xl * b.yl;
    float sum;
    if (((!s.bounds && fp_sem == DME_FP_MMX) || fp_sem == FP_ONE_ROUNDING) && !s.eighth) {
        int isad = 0;
        for (int dy = 0; dy < b.yl; ++dy)
            for (int dx = 0; dx < b.xl; ++dx)
                isad += std::abs(up_value(up, s, dy, dx) - pic.at(b.yp + dy, b.xp + dx));
        sum = (float)isad + start_val;
    } else {
        sum = start_val;
        for (int dy = 0; dy < b.yl; ++dy)
            for (int dx = 0; dx < b.xl; ++dx)
                sum += (float)std::abs(up_value(up, s, dy, dx) - pic.at(b.yp + dy, b.xp + dx));
    }
    if (sum >= best.total) return false;
    best_mv = mv;
    best.total = sum;
    best.mvcost = mvcost;
    best.SAD = sum - start_val;
    return true;
}

// Plain Diff without a start cost (me_utils.cpp:410-479, 573-684): exact
// integer SAD returned as float in both builds.
float up_diff_plain(const Plane &pic, const Plane &up, const Blk &b, const MV &mv, int step,
                    double &terms) {
    if (b.empty()) return 0.f;
    const UpSample s = up_start(b, mv, step, up.w, up.h);
    terms += (double)b.xl * b.yl;
    int isad = 0;
    for (int dy = 0; dy < b.yl; ++dy)
        for (int dx = 0; dx < b.xl; ++dx)
            isad += std::abs(up_value(up, s, dy, dx) - pic.at(b.yp + dy, b.xp + dx));
    return (float)isad;
}

// K5: BiBlockHalfPel / BiBlockQuarterPel::Diff (me_utils.cpp:1172-1543; MMX
// me_utils_mmx.cpp:540-1048): sum |((2*pic - r1) - r2) >> 1| in short
// intermediates, returned as float of the integer sum.  BiBlockEighthPel
// (me_utils.cpp:1545-1785) is the same sum of eighth-pel samples (its float
// adds of integers are exact), except that when the second reference takes
// the bounds-checked path it is read from the first reference's start with
// its own weights (the loop starts at ref_start1, :1764-1768).
float bi_diff(const Plane &pic, const Plane &up1, const Plane &up2, const Blk &b, const MV &m1,
              const MV &m2, int step, double &terms) {
    if (b.empty()) return 0.f;
    const UpSample s1 = up_start(b, m1, step, up1.w, up1.h);
    UpSample s2 = up_start(b, m2, step, up2.w, up2.h);
    if (step == 3 && s2.bounds) {
        s2.rsx = s1.rsx;
        s2.rsy = s1.rsy;
    }
    terms += 2.0 * b.xl * b.yl;
    int sum = 0;
    for (int dy = 0; dy < b.yl; ++dy)
        for (int dx = 0; dx < b.xl; ++dx) {
            const int16_t diff = (int16_t)((pic.at(b.yp + dy, b.xp + dx) << 1) - up_value(up1, s1, dy, dx));
            const int16_t temp = (int16_t)((diff - up_value(up2, s2, dy, dx)) >> 1);
            sum += std::abs((int)temp);
        }
    return (float)sum;
}

// K6: IntraBlockDiff::CalcDC / Diff (me_utils.cpp:284-408; MMX
// simple_intra_block_diff_mmx_4, me_utils_mmx.cpp:111-194): DC is the
// truncating integer mean, cost the SAD against it.
int16_t block_dc(const Plane &pic, const Blk &b) {
    if (b.empty()) return 0;
    int s = 0;
    for (int j = b.yp; j < b.yend(); ++j)
        for (int i = b.xp; i < b.xend(); ++i) s += pic.at(j, i);
    s /= (b.xl * b.yl);
    return (int16_t)s;
}

float intra_diff(const Plane &pic, const Blk &b, int16_t &dc, double &terms) {
    if (b.empty()) { dc = 0; return 0.f; }
    dc = block_dc(pic, b);
    terms += 2.0 * b.xl * b.yl;
    int s = 0;
    for (int j = b.yp; j < b.yend(); ++j)
        for (int i = b.xp; i < b.xend(); ++i) s += std::abs(pic.at(j, i) - dc);
    return (float)s;
}

// ---------------------------------------------------------------------------
// Vector predictors (motion.cpp:488-681)
// ---------------------------------------------------------------------------

MV mv_median3(const MV &a, const MV &b, const MV &c) {
    MV r;
    r.x = a.x + b.x + c.x - std::max(std::max(a.x, b.x), c.x) - std::min(std::min(a.x, b.x), c.x);
    r.y = a.y + b.y + c.y - std::max(std::max(a.y, b.y), c.y) - std::min(std::min(a.y, b.y), c.y);
    return r;
}

MV mv_mean(const MV &a, const MV &b) {
    return MV((a.x + b.x + 1) >> 1, (a.y + b.y + 1) >> 1);
}

MV mv_median_list(const MV *v, int n) {
    if (n == 0) return MV(0, 0);
    if (n == 1) return v[0];
    if (n == 2) return mv_mean(v[0], v[1]);
    if (n == 3) return mv_median3(v[0], v[1], v[2]);
    // n == 4: (sum - max - min + 1) >> 1
    MV s(0, 0), mx = v[0], mn = v[0];
    for (int i = 0; i < 4; ++i) {
        s.x += v[i].x; s.y += v[i].y;
        mx.x = std::max(mx.x, v[i].x); mx.y = std::max(mx.y, v[i].y);
        mn.x = std::min(mn.x, v[i].x); mn.y = std::min(mn.y, v[i].y);
    }
    s.x -= mx.x + mn.x; s.y -= mx.y + mn.y;
    return MV((s.x + 1) >> 1, (s.y + 1) >> 1);
}

// ---------------------------------------------------------------------------
// MEData container (motion.cpp:55-126): zero-initialised arrays.
// ---------------------------------------------------------------------------

struct MEData {
    int xnb = 0, ynb = 0, xsb = 0, ysb = 0, nrefs = 0;
    std::vector<MV> vec[2];
    std::vector<Cost> pc[2];
    std::vector<int> mode;
    std::vector<float> intra;
    std::vector<Cost> bi;
    std::vector<int16_t> dc[3];
    std::vector<int> split;
    std::vector<float> sbc;
    std::vector<float> lam;
    void init(int xn, int yn, int xs, int ys, int nr) {
        xnb = xn; ynb = yn; xsb = xs; ysb = ys; nrefs = nr;
        const size_t B = (size_t)xn * yn, S = (size_t)xs * ys;
        for (int r = 0; r < 2; ++r) { vec[r].assign(B, MV()); pc[r].assign(B, Cost()); }
        mode.assign(B, 0); intra.assign(B, 0.f); bi.assign(B, Cost());
        for (int c = 0; c < 3; ++c) dc[c].assign(B, 0);
        split.assign(S, 0); sbc.assign(S, 0.f); lam.assign(B, 0.f);
    }
    size_t idx(int y, int x) const { return (size_t)y * xnb + x; }
};

// ---------------------------------------------------------------------------
// PixelMatcher (pixel_match.cpp:59-380) with BlockMatcher::FindBestMatchPel
// (block_match.cpp:242-275)
// ---------------------------------------------------------------------------

struct PelLevel {
    int xnb, ynb;
    std::vector<MV> vec;
    std::vector<Cost> pc;
};

// MatchPic at one level against one reference (pixel_match.cpp:242-348).
// full_xr >= 0 selects full search (:249-264, 365-370): the zero list has
// radius (full_xr, full_yr) and the prediction list radius 1.
void match_pic_pel(const Plane &pic, const Plane &ref, const OLB &bp, int level,
                   const PelLevel *guide, PelLevel &out, double &terms, int full_xr = -1, int full_yr = -1) {
    const int r = std::min(level + 1, 5);
    const int pr = full_xr >= 0 ? 1 : r;
    const int X = out.xnb, Y = out.ynb;
    out.vec.assign((size_t)X * Y, MV(0, 0));
    out.pc.assign((size_t)X * Y, Cost());
    for (auto &c : out.pc) c.total = 10000000.0f;

    std::vector<MV> zero_list, guide_list, pred_list;
    if (full_xr >= 0) window(MV(0, 0), full_xr, full_yr, zero_list);
    else window(MV(0, 0), r, r, zero_list);

    auto V = [&](int y, int x) -> MV & { return out.vec[(size_t)y * X + x]; };

    auto do_block = [&](int x, int y, const MV &pred) {
        const Blk b = block_limits(bp, pic.w, pic.h, x, y);
        // lists: zero window, guide window (if any), prediction window
        std::vector<const std::vector<MV> *> lists;
        lists.push_back(&zero_list);
        if (guide) {
            const int xd = bchk(x >> 1, guide->xnb), yd = bchk(y >> 1, guide->ynb);
            const MV g = guide->vec[(size_t)yd * guide->xnb + xd];
            window(MV(g.x * 2, g.y * 2), r, r, guide_list);
            lists.push_back(&guide_list);
        }
        window(pred, pr, pr, pred_list);
        lists.push_back(&pred_list);

        Cost &pc = out.pc[(size_t)y * X + x];
        float best = pc.total;
        MV best_mv = V(y, x);
        if (!b.empty()) {
            std::unordered_set<long long> seen;
            for (const auto *l : lists)
                for (const MV &c : *l) {
                    // AddVect's skip of duplicates (block_match.cpp:170-193)
                    if (!seen.insert(((long long)c.x << 32) ^ (unsigned)c.y).second) continue;
                    terms += (double)b.xl * b.yl;
                    const int s = pel_sad(pic, ref, b, c);
                    if ((float)s < best) { best = (float)s; best_mv = c; }
                }
        }
        V(y, x) = best_mv;
        pc.SAD = best;
        const MV d(best_mv.x - pred.x, best_mv.y - pred.y);
        pc.mvcost = (float)(int16_t)norm1(d);
        pc.total = pc.SAD + 0.0f * pc.mvcost;  // SetTotal(0.0)
    };

    // raster order with the edge prediction rules (pixel_match.cpp:314-347)
    do_block(0, 0, MV(0, 0));
    for (int x = 1; x < X; ++x) do_block(x, 0, V(0, x - 1));
    for (int y = 1; y < Y; ++y) {
        do_block(0, y, V(y - 1, 0));
        for (int x = 1; x < X - 1; ++x)
            do_block(x, y, mv_median3(V(y, x - 1), V(y - 1, x), V(y - 1, x + 1)));
        do_block(X - 1, y, mv_mean(V(y - 1, X - 1), V(y, X - 2)));
    }
}

// ---------------------------------------------------------------------------
// MEData::SetLambdaMap / FindTransitions (motion.cpp:128-171, 201-333).
// The SAD-based transition map computed first is overwritten in full by the
// vector-variation map (motion.cpp:247-309), so only the latter is kept.
// Statistics in x87 long double as in the reference.
// ---------------------------------------------------------------------------

void find_transitions(int X, int Y, const MV *mv, std::vector<char> &trans) {
    std::vector<double> val((size_t)X * Y, 0.0);
    long double total_cost = 0.0;
    for (int j = 1; j < Y - 1; ++j)
        for (int i = 1; i < X - 1; ++i) {
            double v = 0.0;
            const MV &c = mv[(size_t)j * X + i];
            for (int q = -1; q <= 1; ++q)
                for (int p = -1; p <= 1; ++p) {
                    const MV &n = mv[(size_t)(j + q) * X + i + p];
                    v = std::max(v, (double)norm1(MV(n.x - c.x, n.y - c.y)));
                }
            val[(size_t)j * X + i] = v;
            total_cost += v;
        }
    const long double mean_cost = total_cost / static_cast<long double>(X * Y);
    long double sd_cost = 0.0;
    for (int j = 1; j < Y - 1; ++j)
        for (int i = 1; i < X - 1; ++i) {
            double diff = val[(size_t)j * X + i] - mean_cost;
            diff *= diff;
            sd_cost += diff;
        }
    sd_cost /= static_cast<long double>(X * Y);
    sd_cost = std::sqrt(sd_cost);
    const double threshold = mean_cost + 3 * sd_cost;
    trans.assign((size_t)X * Y, 0);
    for (size_t k = 0; k < trans.size(); ++k) trans[k] = val[k] >= threshold;
    // OR-pool 4x4 groups (motion.cpp:311-331)
    for (int j = 0; j < Y / 4; ++j)
        for (int i = 0; i < X / 4; ++i) {
            bool any = false;
            for (int q = 4 * j; q < 4 * (j + 1); ++q)
                for (int p = 4 * i; p < 4 * (i + 1); ++p) any = any || trans[(size_t)q * X + p];
            for (int q = 4 * j; q < 4 * (j + 1); ++q)
                for (int p = 4 * i; p < 4 * (i + 1); ++p) trans[(size_t)q * X + p] = any;
        }
}

void lambda_map(int X, int Y, int nrefs, const MV *v1, const MV *v2, float lambda, float *out) {
    std::vector<char> t1, t2;
    find_transitions(X, Y, v1, t1);
    if (nrefs > 1) find_transitions(X, Y, v2, t2);
    for (int j = 0; j < Y; ++j)
        for (int i = 0; i < X; ++i) {
            const size_t k = (size_t)j * X + i;
            float m;
            if (nrefs == 1) {
                m = t1[k] ? 0.0f : lambda;
            } else {
                if (t1[k] && t2[k]) m = 0.0f;
                else if (t1[k] || t2[k]) m = (float)(lambda / 4.0);
                else m = lambda;
            }
            if (i < 4 || j < 4) m = (float)(m / 5.0);
            out[k] = m;
        }
}

// SetLambdaMap(level, map, wt) (motion.cpp:173-199): max over the
// 2^(2-level) square group, times wt.
void level_lambda(const MEData &base, int level, float wt, int xn, int yn, std::vector<float> &out) {
    const int f = 1 << (2 - level);
    out.assign((size_t)xn * yn, 0.f);
    for (int j = 0; j < yn; ++j)
        for (int i = 0; i < xn; ++i) {
            float m = base.lam[base.idx(f * j, f * i)];
            for (int q = f * j; q < f * (j + 1); ++q)
                for (int p = f * i; p < f * (i + 1); ++p) m = std::max(base.lam[base.idx(q, p)], m);
            m *= wt;
            out[(size_t)j * xn + i] = m;
        }
}

// ---------------------------------------------------------------------------
// SubpelRefine (me_subpel.cpp:65-176) with BlockMatcher::RefineMatchSubp
// (block_match.cpp:327-467)
// ---------------------------------------------------------------------------

// pred_from non-NULL (the fast form): the predictor of every block is taken
// from that field (its integer vectors scaled to 1/2^prec pel) instead of the
// vectors refined so far, so that every block is independent.
void subpel_ref(const Plane &pic, const Plane &up, const Geo &g, int prec, int fp_sem, MEData &me,
                int r, double &terms, const std::vector<MV> *pred_from = nullptr) {
    const int X = me.xnb, Y = me.ynb;
    std::vector<MV> &V = me.vec[r];
    const std::vector<MV> &PV = pred_from ? *pred_from : V;
    std::vector<Cost> &C = me.pc[r];
    const OLB &bp = g.l[2];
    // neighbour order of m_nshift (me_subpel.cpp:46-63)
    const int nsx[4] = {-1, -1, 0, 1}, nsy[4] = {0, -1, -1, -1};
    for (int y = 0; y < Y; ++y)
        for (int x = 0; x < X; ++x) {
            MV nb[4];
            int n = 0;
            for (int k = 0; k < 4; ++k) {
                const int xx = x + nsx[k], yy = y + nsy[k];
                if (xx >= 0 && yy >= 0 && xx < X && yy < Y) nb[n++] = PV[me.idx(yy, xx)];
            }
            const MV pred = mv_median_list(nb, n);
            const float lambda = me.lam[me.idx(y, x)];
            const size_t k = me.idx(y, x);
            const Blk b = block_limits(bp, pic.w, pic.h, x, y);
            const MV pel = V[k];
            auto getvarup = [&](const MV &m) {
                const int a = norm1(MV(m.x - pred.x, m.y - pred.y));
                return (float)(int16_t)std::min(a, norm1(m));
            };
            C[k].mvcost = getvarup(MV(pel.x << prec, pel.y << prec));
            C[k].total = C[k].SAD + lambda * C[k].mvcost;
            Cost best = C[k];
            MV best_mv = pel;
            const float lim = (float)(2 * b.xl * b.yl);
            if (best.SAD < lim) {
                V[k] = MV(pel.x << prec, pel.y << prec);
                continue;
            }
            Cost pcost;
            pcost.mvcost = 0;
            pcost.SAD = up_diff_plain(pic, up, b, pred, prec, terms);
            pcost.total = pcost.SAD;
            if (pcost.SAD < lim) {
                V[k] = pred;
                C[k] = pcost;
                continue;
            }
            bool bailed = false;
            for (int i = 1; i <= prec; ++i) {
                best_mv = MV(best_mv.x << 1, best_mv.y << 1);
                MV tb = best_mv;
                auto test = [&](const MV &c) {
                    up_diff_update(pic, up, b, c, i, getvarup(MV(c.x << (prec - i), c.y << (prec - i))),
                                   lambda, fp_sem, best, tb, terms);
                };
                test(MV(best_mv.x - 1, best_mv.y));
                test(MV(best_mv.x + 1, best_mv.y));
                test(MV(best_mv.x, best_mv.y - 1));
                test(MV(best_mv.x, best_mv.y + 1));
                if (tb.x != best_mv.x) {
                    const MV nb2 = tb;
                    test(MV(nb2.x, nb2.y - 1));
                    test(MV(nb2.x, nb2.y + 1));
                } else if (tb.y != best_mv.y) {
                    const MV nb2 = tb;
                    test(MV(nb2.x - 1, nb2.y));
                    test(MV(nb2.x + 1, nb2.y));
                }
                best_mv = tb;
                if ((double)best.total > 1.1 * (double)pcost.total) {
                    V[k] = pred;
                    C[k] = pcost;
                    bailed = true;
                    break;
                }
            }
            if (bailed) continue;
            V[k] = best_mv;
            C[k] = best;
        }
}

// ---------------------------------------------------------------------------
// ModeDecider (me_mode_decn.cpp:66-573)
// ---------------------------------------------------------------------------

struct ModeCtx {
    const Geo *g;
    const Plane *pic;
    const Plane *up[2];
    int nrefs, step, fp_sem;
    MEData *lv[3];  // level 0, 1 (temporary) and 2 (the picture's)
    double *terms;
    // the fast form: level-1/0 predictors from this (pre-merge) field instead
    // of the level-2 vectors merged so far
    const std::vector<MV> *pred_v2[2] = {nullptr, nullptr};
};

// DoME (me_mode_decn.cpp:279-364) with FindBestMatchSubp
// (block_match.cpp:277-325) for one reference.
void mode_do_me_ref(ModeCtx &m, int xpos, int ypos, int level, int r) {
    MEData &me = *m.lv[level];
    const MEData &gd = *m.lv[level + 1];
    const MEData &b2 = *m.lv[2];
    const int gx = xpos << 1, gy = ypos << 1;
    const int xb = xpos << (2 - level), yb = ypos << (2 - level);
    const float lambda = me.lam[me.idx(ypos, xpos)];
    MV cands[4];
    int nc = 0;
    for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
            const MV c = gd.vec[r][gd.idx(gy + j, gx + i)];
            bool dup = false;
            for (int q = 0; q < nc; ++q) dup = dup || cands[q] == c;
            if (!dup) cands[nc++] = c;
        }
    MV pred;
    const std::vector<MV> &V2 = m.pred_v2[r] ? *m.pred_v2[r] : b2.vec[r];
    if (xb > 0 && yb > 0)
        pred = mv_median3(V2[b2.idx(yb, xb - 1)], V2[b2.idx(yb - 1, xb - 1)], V2[b2.idx(yb - 1, xb)]);
    else if (xb == 0 && yb > 0)
        pred = mv_mean(V2[b2.idx(yb - 1, xb)], V2[b2.idx(yb - 1, xb + 1)]);
    else if (xb > 0 && yb == 0)
        pred = mv_mean(V2[b2.idx(yb, xb - 1)], V2[b2.idx(yb + 1, xb - 1)]);
    else
        pred = MV(0, 0);
    const size_t k = me.idx(ypos, xpos);
    me.pc[r][k].total = 100000000.0f;
    const Blk b = block_limits(m.g->l[level], m.pic->w, m.pic->h, xpos, ypos);
    Cost best = me.pc[r][k];
    best.total = 100000000.0f;
    MV best_mv = me.vec[r][k];
    for (int q = 0; q < nc; ++q) {
        const MV &c = cands[q];
        const int a = norm1(MV(c.x - pred.x, c.y - pred.y));
        const float mvcost = (float)(int16_t)std::min(a, norm1(c));
        up_diff_update(*m.pic, *m.up[r], b, c, m.step, mvcost, lambda, m.fp_sem, best, best_mv, *m.terms);
    }
    me.vec[r][k] = best_mv;
    me.pc[r][k] = best;
}

// DoUnitDecn (me_mode_decn.cpp:368-462)
float mode_unit(ModeCtx &m, int xpos, int ypos, int level) {
    MEData &me = *m.lv[level];
    const size_t k = me.idx(ypos, xpos);
    const float loc_lambda = me.lam[k];
    const float mode_cost = 0.0f;
    const float lf = m.g->level_factor[level];
    const Blk b = block_limits(m.g->l[level], m.pic->w, m.pic->h, xpos, ypos);
    me.mode[k] = DME_REF1_ONLY;
    me.pc[0][k].total *= lf;
    float min_unit_cost = me.pc[0][k].total + mode_cost;
    float best_SAD = me.pc[0][k].SAD;
    if (m.nrefs > 1) {
        me.pc[1][k].total *= lf;
        float unit_cost = me.pc[1][k].total + mode_cost;
        if (unit_cost < min_unit_cost) {
            me.mode[k] = DME_REF2_ONLY;
            min_unit_cost = unit_cost;
            best_SAD = me.pc[1][k].SAD;
        }
        Cost &bc = me.bi[k];
        bc.mvcost = me.pc[0][k].mvcost + me.pc[1][k].mvcost;
        bc.SAD = bi_diff(*m.pic, *m.up[0], *m.up[1], b, me.vec[0][k], me.vec[1][k], m.step, *m.terms);
        bc.total = bc.SAD + loc_lambda * bc.mvcost;
        bc.total *= lf;
        unit_cost = bc.total + mode_cost;
        if (unit_cost < min_unit_cost) {
            me.mode[k] = DME_REF1AND2;
            min_unit_cost = unit_cost;
            best_SAD = bc.SAD;
        }
    }
    const OLB &bp = m.g->l[level];
    if (level == 2 && best_SAD > 4.0 * bp.xblen * bp.yblen) {
        me.intra[k] = intra_diff(*m.pic, b, me.dc[0][k], *m.terms);
        me.intra[k] *= lf;
        const float unit_cost = me.intra[k] + mode_cost;
        if (unit_cost < min_unit_cost && me.intra[k] < 0.85 * best_SAD) {
            me.mode[k] = DME_INTRA;
            min_unit_cost = unit_cost;
        }
    }
    return min_unit_cost;
}

// DoLevelDecn (me_mode_decn.cpp:199-276)
void mode_level(ModeCtx &m, int xsb, int ysb, int level) {
    MEData &b2 = *m.lv[2];
    const int xs = xsb << level, ys = ysb << level;
    const int xe = xs + (1 << level), ye = ys + (1 << level);
    float sb_cost = 0.0f;
    for (int j = ys; j < ye; ++j)
        for (int i = xs; i < xe; ++i) {
            if (level < 2) {
                mode_do_me_ref(m, i, j, level, 0);
                if (m.nrefs > 1) mode_do_me_ref(m, i, j, level, 1);
            }
            sb_cost += mode_unit(m, i, j, level);
        }
    const size_t s = (size_t)ysb * b2.xsb + xsb;
    if (level == 2) {
        b2.split[s] = 2;
        b2.sbc[s] = sb_cost;
    }
    if (level < 2 && sb_cost <= b2.sbc[s]) {
        b2.sbc[s] = sb_cost;
        b2.split[s] = level;
        const MEData &lv = *m.lv[level];
        for (int j = ys; j < ye; ++j)
            for (int i = xs; i < xe; ++i) {
                const size_t u = lv.idx(j, i);
                for (int v = j << (2 - level); v < (j + 1) << (2 - level); ++v)
                    for (int w = i << (2 - level); w < (i + 1) << (2 - level); ++w) {
                        const size_t d = b2.idx(v, w);
                        b2.mode[d] = lv.mode[u];
                        b2.dc[0][d] = lv.dc[0][u];
                        b2.vec[0][d] = lv.vec[0][u];
                        if (m.nrefs > 1) b2.vec[1][d] = lv.vec[1][u];
                    }
            }
    }
}

// SetDC for one component (me_mode_decn.cpp:523-565).  The last row is never
// written (loop bound LengthY()-1).
void set_dc(const Plane &pic, const OLB &bp, MEData &me, int comp) {
    const int X = me.xnb, Y = me.ynb;
    std::vector<int16_t> t((size_t)X * Y);
    for (int y = 0; y < Y; ++y)
        for (int x = 0; x < X; ++x) t[(size_t)y * X + x] = block_dc(pic, block_limits(bp, pic.w, pic.h, x, y));
    std::vector<int16_t> &dc = me.dc[comp];
    auto T = [&](int y, int x) { return (int)t[(size_t)y * X + x]; };
    for (int x = 0; x < X; ++x) dc[x] = t[x];
    for (int y = 1; y < Y - 1; ++y) {
        dc[(size_t)y * X] = t[(size_t)y * X];
        for (int x = 1; x < X - 1; ++x)
            dc[(size_t)y * X + x] = (int16_t)((T(y - 1, x - 1) + 3 * T(y - 1, x) + T(y - 1, x + 1) +
                                               3 * T(y, x - 1) + 3 * T(y, x + 1) + T(y + 1, x - 1) +
                                               3 * T(y + 1, x) + T(y + 1, x + 1) + 8) >> 4);
        dc[(size_t)y * X + X - 1] = t[(size_t)y * X + X - 1];
    }
}

Plane make_plane(const int16_t *p, int w, int h) {
    Plane q(w, h);
    std::memcpy(q.d.data(), p, sizeof(int16_t) * (size_t)w * h);
    return q;
}

// EncPicture::Combine, 4:2:0 branch (enc_picture.cpp:285-367; CombinedData,
// :237-254): each sample of a 2 x 2 luma group with the group's chroma,
// ValueType(sqrt(u^2 + v^2 + (y + 128)^2) - 128), the squares and sums in
// float (exact for 8-bit samples), sqrt in float, the subtraction in double,
// truncated toward zero.
Plane combine(const int16_t *y, const int16_t *u, const int16_t *v, int W, int H) {
    Plane c(W, H);
    const int CW = W >> 1;
    for (int j = 0; j < H; ++j)
        for (int i = 0; i < W; ++i) {
            const int k = (j >> 1) * CW + (i >> 1);
            float val = float(u[k]);
            val *= val;
            float valc = val;
            val = float(v[k]);
            val *= val;
            valc += val;
            float valy = float(y[(size_t)j * W + i]) + 128.0f;
            valy *= valy;
            c.at(j, i) = (int16_t)(std::sqrt(valc + valy) - 128.0);
        }
    return c;
}

void export_medata(const MEData &me, dme_medata *out) {
    if (!out) return;
    const size_t B = (size_t)me.xnb * me.ynb, S = (size_t)me.xsb * me.ysb;
    for (int r = 0; r < 2; ++r) {
        if (out->vectors[r])
            for (size_t k = 0; k < B; ++k) { out->vectors[r][k].x = me.vec[r][k].x; out->vectors[r][k].y = me.vec[r][k].y; }
        if (out->pred_costs[r])
            for (size_t k = 0; k < B; ++k) {
                out->pred_costs[r][k].SAD = me.pc[r][k].SAD;
                out->pred_costs[r][k].mvcost = me.pc[r][k].mvcost;
                out->pred_costs[r][k].total = me.pc[r][k].total;
            }
    }
    if (out->mode) for (size_t k = 0; k < B; ++k) out->mode[k] = me.mode[k];
    if (out->intra_costs) for (size_t k = 0; k < B; ++k) out->intra_costs[k] = me.intra[k];
    if (out->bipred_costs)
        for (size_t k = 0; k < B; ++k) {
            out->bipred_costs[k].SAD = me.bi[k].SAD;
            out->bipred_costs[k].mvcost = me.bi[k].mvcost;
            out->bipred_costs[k].total = me.bi[k].total;
        }
    for (int c = 0; c < 3; ++c)
        if (out->dc[c]) for (size_t k = 0; k < B; ++k) out->dc[c][k] = me.dc[c][k];
    if (out->sb_split) for (size_t k = 0; k < S; ++k) out->sb_split[k] = me.split[k];
    if (out->sb_costs) for (size_t k = 0; k < S; ++k) out->sb_costs[k] = me.sbc[k];
    if (out->lambda_map) for (size_t k = 0; k < B; ++k) out->lambda_map[k] = me.lam[k];
    // PictureCompressor::IntraModeAnalyse (picture_compress.cpp:300-320):
    // INTRA blocks over all blocks, divided in double, kept as a float
    // (MEData::SetIntraBlockRatio, motion.h:359)
    if (out->intra_block_ratio) {
        size_t c = 0;
        for (size_t k = 0; k < B; ++k) c += me.mode[k] == DME_INTRA;
        *out->intra_block_ratio = static_cast<float>(static_cast<double>(c) / static_cast<double>(B));
    }
}

// ---------------------------------------------------------------------------
// Motion compensation (MotionCompensator, mot_comp.cpp), restated in the
// reference's own structure: block rows processed in order into a strip of
// Yblen rows that carries the overlap to the next row (CompensateComponent,
// :238-539), blocks merged horizontally by the superblock split (:321-400),
// per block prediction, reference weights, spatial weights and
// accumulation (CompensateBlock, :541-603).
// ---------------------------------------------------------------------------

// CreateBlock (mot_comp.cpp:735-804): the OBMC window of a block of
// xlen x ylen with separations xbsep, ybsep; FullX / FullY false keep the
// leading edge flat (picture edges).
Plane mc_create_block(int xlen, int ylen, int xbsep, int ybsep, bool fullx, bool fully) {
    std::vector<int> hw(xlen), vw(ylen);
    const int xoff = (xlen - xbsep) / 2, yoff = (ylen - ybsep) / 2;
    if (xoff != 1) {
        for (int x = 0; x < 2 * xoff; ++x) {
            hw[x] = 1 + (6 * x + xoff - 1) / (2 * xoff - 1);
            hw[x + xbsep] = 8 - hw[x];
        }
    } else {
        hw[0] = hw[1 + xbsep] = 3;
        hw[1] = hw[xbsep] = 5;
    }
    for (int x = 2 * xoff; x < xbsep; ++x) hw[x] = 8;
    if (yoff != 1) {
        for (int y = 0; y < 2 * yoff; ++y) {
            vw[y] = 1 + (6 * y + yoff - 1) / (2 * yoff - 1);
            vw[y + ybsep] = 8 - vw[y];
        }
    } else {
        vw[0] = vw[1 + ybsep] = 3;
        vw[1] = vw[ybsep] = 5;
    }
    for (int y = 2 * yoff; y < ybsep; ++y) vw[y] = 8;
    if (!fullx) for (int x = 0; x < 2 * xoff; ++x) hw[x] = 8;
    if (!fully) for (int y = 0; y < 2 * yoff; ++y) vw[y] = 8;
    Plane w(xlen, ylen);
    for (int y = 0; y < ylen; ++y)
        for (int x = 0; x < xlen; ++x) w.at(y, x) = (int16_t)(vw[y] * hw[x]);
    return w;
}

Plane mc_flip_x(const Plane &a) {
    Plane b(a.w, a.h);
    for (int y = 0; y < a.h; ++y)
        for (int x = 0; x < a.w; ++x) b.at(y, x) = a.at(y, a.w - 1 - x);
    return b;
}

Plane mc_flip_y(const Plane &a) {
    Plane b(a.w, a.h);
    for (int y = 0; y < a.h; ++y)
        for (int x = 0; x < a.w; ++x) b.at(y, x) = a.at(a.h - 1 - y, x);
    return b;
}

// CalculateWeights (mot_comp.cpp:703-733): the nine edge variants.
void mc_weights(int xlen, int ylen, int xbsep, int ybsep, Plane w[9]) {
    w[0] = mc_create_block(xlen, ylen, xbsep, ybsep, false, false);
    w[3] = mc_create_block(xlen, ylen, xbsep, ybsep, false, true);
    w[1] = mc_create_block(xlen, ylen, xbsep, ybsep, true, false);
    w[4] = mc_create_block(xlen, ylen, xbsep, ybsep, true, true);
    w[5] = mc_flip_x(w[3]);
    w[2] = mc_flip_x(w[0]);
    w[6] = mc_flip_y(w[0]);
    w[8] = mc_flip_x(w[6]);
    w[7] = mc_flip_y(w[1]);
}

// BlockPixelPred for pixel, half, quarter and eighth precision (mot_comp.cpp:846-1230;
// MMX mot_comp_mmx.cpp:69-354 gives the same values): the block of
// ylen x xlen from the up-converted reference, clamped reads via BChk
// against (2 size - 1).
void mc_block_pred(Plane &blk, int sx, int sy, int prec, const Plane &up, int pw, int ph, MV mv) {
    const int tx = (pw << 1) - 1, ty = (ph << 1) - 1;
    int rx0, ry0, fx = 0, fy = 0;
    if (prec == 0) {
        rx0 = (sx + mv.x) << 1;
        ry0 = (sy + mv.y) << 1;
    } else if (prec == 1) {
        rx0 = (sx << 1) + mv.x;
        ry0 = (sy << 1) + mv.y;
    } else if (prec == 2) {
        rx0 = (sx << 1) + (mv.x >> 1);
        ry0 = (sy << 1) + (mv.y >> 1);
        fx = mv.x & 1;
        fy = mv.y & 1;
    } else {  // eighth pel (:1108-1230): bilinear in quarter steps of the half-pel plane
        rx0 = (sx << 1) + (mv.x >> 2);
        ry0 = (sy << 1) + (mv.y >> 2);
        fx = mv.x & 3;
        fy = mv.y & 3;
    }
    const int one = prec == 3 ? 4 : 2, sh = prec == 3 ? 4 : 2;
    const int w0 = (one - fx) * (one - fy), w1 = fx * (one - fy), w2 = (one - fx) * fy, w3 = fx * fy;
    for (int y = 0; y < blk.h; ++y) {
        const int uy = ry0 + 2 * y;
        for (int x = 0; x < blk.w; ++x) {
            const int ux = rx0 + 2 * x;
            if (prec < 2) {
                blk.at(y, x) = up.at(bchk(uy, ty), bchk(ux, tx));
            } else {
                const int v = w0 * up.at(bchk(uy, ty), bchk(ux, tx)) + w1 * up.at(bchk(uy, ty), bchk(ux + 1, tx)) +
                              w2 * up.at(bchk(uy + 1, ty), bchk(ux, tx)) + w3 * up.at(bchk(uy + 1, ty), bchk(ux + 1, tx));
                blk.at(y, x) = (int16_t)((v + (1 << (sh - 1))) >> sh);
            }
        }
    }
}

struct McArrays {
    int xnb, ynb, xsb, nrefs;
    const dme_mv *v[2];
    const int32_t *mode;
    const int16_t *dc;
    const int32_t *split;
};

// CompensateComponent (mot_comp.cpp:238-539) for one component.
void mc_component(const OLB &bp, int prec, bool mmx, int pw, int ph, int shift, const McArrays &m,
                  const Plane *up[2], int direction, int16_t *out, int lx, int ly) {
    const int bpm = m.xnb / m.xsb, bps = bpm >> 1;  // blocks per superblock / sub-superblock row
    const int xoff = (bp.xblen - bp.xbsep) / 2, yoff = (bp.yblen - bp.ybsep) / 2;
    const int mb_xlen = bp.xblen * bpm - (bp.xblen - bp.xbsep) * (bpm - 1);
    const int mb_xsep = mb_xlen - (bp.xblen - bp.xbsep);
    const int sb_xlen = bp.xblen * bps - (bp.xblen - bp.xbsep) * (bps - 1);
    const int sb_xsep = sb_xlen - (bp.xblen - bp.xbsep);
    Plane wb[9], wm[9], ws[9];
    mc_weights(bp.xblen, bp.yblen, bp.xbsep, bp.ybsep, wb);
    mc_weights(mb_xlen, bp.yblen, mb_xsep, bp.ybsep, wm);
    mc_weights(sb_xlen, bp.yblen, sb_xsep, bp.ybsep, ws);
    // NUM_USED_BLKS (mot_comp.cpp:54)
    const int ynb = std::min((ph + bp.ybsep + (bp.yblen - bp.ybsep) / 2 - 1) / bp.ybsep, m.ynb);
    const int xnb = std::min((pw + bp.xbsep + (bp.xblen - bp.xbsep) / 2 - 1) / bp.xbsep, m.xnb);
    Plane strip(lx, bp.yblen);  // pic_data: one row of blocks
    int save_from = bp.ybsep - yoff;
    const bool overlap = bp.yblen - bp.ybsep > 0;
    const int x_end = std::min(lx, pw), y_end = std::min(ly, ph);
    int py = -yoff;
    for (int yb = 0; yb < ynb; ++yb) {
        int px = -xoff, inc = 0;
        for (int xb = 0; xb < xnb; xb += inc) {
            const int split = m.split[(size_t)(yb / bpm) * m.xsb + xb / bpm];
            const int bx = split == 0 ? bpm : (split == 1 ? bps : 1);
            int wi;
            const int row = py < 0 ? 0 : (yb + 1 < ynb ? 1 : 2);
            if (px >= 0 && xb + bx < xnb) wi = 3 * row + 1;
            else if (px < 0) wi = 3 * row;
            else wi = 3 * row + 2;
            const Plane &wt = split == 0 ? wm[wi] : (split == 1 ? ws[wi] : wb[wi]);
            inc = bx;
            const size_t k = (size_t)yb * m.xnb + xb;
            const int mode = m.mode[k];
            MV mv1(m.v[0][k].x >> shift, m.v[0][k].y >> shift);
            const dme_mv *v2 = m.nrefs > 1 ? m.v[1] : m.v[0];
            MV mv2(v2[k].x >> shift, v2[k].y >> shift);
            // CompensateBlock (mot_comp.cpp:541-603)
            const int sx = std::max(px, 0), sy = std::max(py, 0);
            const int ex = std::min(px + wt.w, pw), ey = std::min(py + wt.h, ph);
            if (sx < ex && sy < ey) {
                Plane v1(ex - sx, ey - sy), v2b(ex - sx, ey - sy);
                if (mode == DME_REF1_ONLY) mc_block_pred(v1, sx, sy, prec, *up[0], pw, ph, mv1);
                else if (mode == DME_REF2_ONLY) mc_block_pred(v1, sx, sy, prec, *up[1], pw, ph, mv2);
                else if (mode == DME_REF1AND2) {
                    mc_block_pred(v1, sx, sy, prec, *up[0], pw, ph, mv1);
                    mc_block_pred(v2b, sx, sy, prec, *up[1], pw, ph, mv2);
                } else {
                    for (auto &e : v1.d) e = m.dc[k];  // DCBlock (:605-618)
                }
                // AdjustBlockByRefWeights, default weights (:620-683)
                if (mode == DME_REF1AND2)
                    for (size_t i = 0; i < v1.d.size(); ++i) v1.d[i] = (int16_t)((v1.d[i] + v2b.d[i] + 1) >> 1);
                // AdjustBlockBySpatialWeights (:685-701) and accumulation
                for (int y = 0; y < v1.h; ++y)
                    for (int x = 0; x < v1.w; ++x) {
                        const int16_t val = (int16_t)(v1.at(y, x) * wt.at(sy - py + y, sx - px + x));
                        strip.at(y, sx + x) = (int16_t)(strip.at(y, sx + x) + val);
                    }
            }
            px += bp.xbsep * inc;
        }
        if (direction == DME_MC_SUBTRACT) {
            const int s0 = std::max(0, py);
            int e0 = std::min(py + bp.ybsep, y_end);
            if (yb == ynb - 1) e0 = std::min(ly, y_end);
            for (int i = s0, r = 0; i < e0; ++i, ++r) {
                int16_t *o = out + (size_t)i * lx;
                for (int j = 0; j < x_end; ++j) o[j] = (int16_t)(o[j] - (int16_t)((strip.at(r, j) + 32) >> 6));
                for (int j = pw; j < lx; ++j) o[j] = 0;
            }
        } else {
            const int s0 = std::max(0, py);
            int e0 = std::min(py + bp.ybsep, ly);
            if (yb == ynb - 1) {
                e0 += bp.yblen - bp.ybsep;
                if (e0 > ph) e0 = ph;
            }
            for (int i = s0, r = 0; i < e0; ++i, ++r) {
                int16_t *o = out + (size_t)i * lx;
                // The MMX build adds the rounding offset in 16-bit lanes over the
                // first (width & ~3) samples (mot_comp_mmx.cpp:408-424), the C
                // build in int (mot_comp.cpp:474-482).
                for (int j = 0; j < pw; ++j) {
                    const int sum = strip.at(r, j) + 32;
                    const int16_t p = (int16_t)(mmx && j < (pw & ~3) ? (int16_t)sum >> 6 : sum >> 6);
                    o[j] = (int16_t)(o[j] + p);
                }
                for (int j = pw; j < lx; ++j) o[j] = o[pw - 1];
            }
        }
        py += bp.ybsep;
        if (overlap) {
            const int keep = bp.yblen - save_from;
            for (int r = 0; r < keep; ++r)
                for (int x = 0; x < lx; ++x) strip.at(r, x) = strip.at(r + save_from, x);
            for (int r = keep; r < bp.yblen; ++r)
                for (int x = 0; x < lx; ++x) strip.at(r, x) = 0;
            save_from = bp.ybsep;
        } else {
            std::fill(strip.d.begin(), strip.d.end(), (int16_t)0);
        }
    }
    if (direction == DME_MC_SUBTRACT) {
        for (int y = ph; y < ly; ++y)
            for (int x = 0; x < lx; ++x) out[(size_t)y * lx + x] = 0;
    } else {
        for (int y = ph; y < ly; ++y)
            for (int x = 0; x < lx; ++x) out[(size_t)y * lx + x] = out[(size_t)(ph - 1) * lx + x];
    }
}


// ---------------------------------------------------------------------------
// The fast form of the learned path (no counterpart in the reference; the
// build's own algorithm, restated here so that its GPU kernels are pinned
// bit-exactly given the same level-0 integer vectors).  Sub-pel refinement
// and mode decision run with predictors that do not depend on decisions
// made in the same pass (the integer vectors for the sub-pel search, the
// pre-merge level-2 vectors for the level-1/0 units), so every block and
// superblock is independent.  Then the reference's Lagrangian is evaluated
// on the final vectors with the reference's own predictors:
//  - level-2 block costs (the PredCosts SubpelRefine leaves,
//    block_match.cpp:327-467): SAD of the final vector, mvcost =
//    GetVarUp(GetPred of the final field, mv) (me_subpel.cpp:145-176), total
//    = SAD + lambda mvcost (SetTotal, motion.h:201); a block with no samples
//    costs {0, 0, 0} (the reference's result for it, whatever its vector);
//  - a superblock split into level-1 or level-0 units: each unit's costs
//    with the predictor DoME forms (me_mode_decn.cpp:308-321) from the
//    level-2 vectors as the raster pass of DoModeDecn sees them: the final
//    (merged) vectors of earlier superblocks, the pre-merge ones inside the
//    superblock; then DoUnitDecn (me_mode_decn.cpp:368-462) and the sum.
// ---------------------------------------------------------------------------

// GetPred (me_subpel.cpp:145-176) at block (x, y) of field V.
MV subpel_pred(const std::vector<MV> &V, int X, int Y, int x, int y) {
    const int nsx[4] = {-1, -1, 0, 1}, nsy[4] = {0, -1, -1, -1};
    MV nb[4];
    int n = 0;
    for (int k = 0; k < 4; ++k) {
        const int xx = x + nsx[k], yy = y + nsy[k];
        if (xx >= 0 && yy >= 0 && xx < X && yy < Y) nb[n++] = V[(size_t)yy * X + xx];
    }
    return mv_median_list(nb, n);
}

float getvarup(const MV &pred, const MV &m) {
    const int a = norm1(MV(m.x - pred.x, m.y - pred.y));
    return (float)(int16_t)std::min(a, norm1(m));
}

// Exact level-2 costs of the final sub-pel vectors of reference r.  A final
// vector equal to the scaled integer vector keeps the pel search's SAD
// (RefineMatchSubp starts from it and no candidate replaced it), whose reads
// clamp in the picture rather than in the up-converted plane; v0/pel_sad are
// the integer vectors and their SADs.
void fast_l2_costs(const Plane &pic, const Plane &up, const Geo &g, int prec, MEData &me, int r,
                   const std::vector<MV> &v0, const std::vector<float> &pel_sad, double &terms) {
    const int X = me.xnb, Y = me.ynb;
    for (int y = 0; y < Y; ++y)
        for (int x = 0; x < X; ++x) {
            const size_t k = me.idx(y, x);
            const Blk b = block_limits(g.l[2], pic.w, pic.h, x, y);
            Cost c;
            if (b.empty()) {
                c.SAD = c.mvcost = c.total = 0.f;
            } else {
                const MV mv = me.vec[r][k];
                const bool at_pel = mv == MV(v0[k].x << prec, v0[k].y << prec);
                c.SAD = at_pel ? pel_sad[k] : up_diff_plain(pic, up, b, mv, prec, terms);
                c.mvcost = getvarup(subpel_pred(me.vec[r], X, Y, x, y), mv);
                c.total = c.SAD + me.lam[k] * c.mvcost;
            }
            me.pc[r][k] = c;
        }
}

// The exact cost of superblock (xs, ys) at its split level s (0 or 1), from
// the unit vectors of level s, and the unit modes that go with it.
void fast_eval_sb(ModeCtx &m, const std::vector<MV> *V2, int xs, int ys) {
    MEData &b2 = *m.lv[2];
    const size_t si = (size_t)ys * b2.xsb + xs;
    const int s = b2.split[si];
    if (s == 2) return;
    const MEData &lv = *m.lv[s];
    const float lf = m.g->level_factor[s];
    // the level-2 vector field as DoModeDecn's raster pass sees it at (xs, ys)
    auto field = [&](int r, int u, int v) -> MV {
        const int sx = u >> 2, sy = v >> 2;
        const bool earlier = sy < ys || (sy == ys && sx < xs);
        return earlier ? b2.vec[r][b2.idx(v, u)] : V2[r][b2.idx(v, u)];
    };
    float sb_cost = 0.0f;
    for (int j = ys << s; j < (ys + 1) << s; ++j)
        for (int i = xs << s; i < (xs + 1) << s; ++i) {
            const size_t k = lv.idx(j, i);
            const int xb = i << (2 - s), yb = j << (2 - s);
            const float lambda = lv.lam[k];
            const Blk b = block_limits(m.g->l[s], m.pic->w, m.pic->h, i, j);
            // the unit's vectors: a merge copied them to all its blocks
            MV uv[2];
            for (int r = 0; r < m.nrefs; ++r) uv[r] = b2.vec[r][b2.idx(yb, xb)];
            Cost c[2];
            for (int r = 0; r < m.nrefs; ++r) {
                MV pred;
                if (xb > 0 && yb > 0)
                    pred = mv_median3(field(r, xb - 1, yb), field(r, xb - 1, yb - 1), field(r, xb, yb - 1));
                else if (xb == 0 && yb > 0)
                    pred = mv_mean(field(r, xb, yb - 1), field(r, xb + 1, yb - 1));
                else if (xb > 0 && yb == 0)
                    pred = mv_mean(field(r, xb - 1, yb), field(r, xb - 1, yb + 1));
                else
                    pred = MV(0, 0);
                const MV v = uv[r];
                c[r].SAD = c[r].mvcost = 0.f;
                c[r].total = 100000000.0f;
                MV dummy = v;
                // (a unit with no samples keeps {0, 0, 1e8}: FindBestMatchSubp's
                // start, from zero-initialised level-1/0 MEData)
                up_diff_update(*m.pic, *m.up[r], b, v, m.step, getvarup(pred, v), lambda, DME_FP_MMX, c[r], dummy,
                               *m.terms);
            }
            // DoUnitDecn at level s (no intra below level 2)
            int mode = DME_REF1_ONLY;
            float best = c[0].total * lf;
            if (m.nrefs > 1) {
                const float t2 = c[1].total * lf;
                if (t2 < best) { mode = DME_REF2_ONLY; best = t2; }
                Cost bc;
                bc.mvcost = c[0].mvcost + c[1].mvcost;
                bc.SAD = bi_diff(*m.pic, *m.up[0], *m.up[1], b, uv[0], uv[1], m.step, *m.terms);
                bc.total = bc.SAD + lambda * bc.mvcost;
                bc.total *= lf;
                if (bc.total < best) { mode = DME_REF1AND2; best = bc.total; }
            }
            sb_cost += best;
            for (int v = yb; v < (j + 1) << (2 - s); ++v)
                for (int u = xb; u < (i + 1) << (2 - s); ++u) b2.mode[b2.idx(v, u)] = mode;
        }
    b2.sbc[si] = sb_cost;
}

}  // namespace

// ===========================================================================
// C entry points
// ===========================================================================

extern "C" const char *oracle_last_error(void) { return g_err.c_str(); }

extern "C" int oracle_geometry(const dme_params *p, int width, int height, dme_geometry *out) {
    Geo g;
    int rc = make_geo(p, width, height, g);
    if (rc) return rc;
    if (!out) return DME_OK;
    std::memset(out, 0, sizeof(*out));
    out->width = g.W; out->height = g.H; out->chroma_width = g.CW; out->chroma_height = g.CH;
    out->xnum_sb = g.xsb; out->ynum_sb = g.ysb; out->xnum_blocks = g.xnb; out->ynum_blocks = g.ynb;
    out->depth = g.depth;
    for (int i = 0; i <= g.depth; ++i) {
        out->level_width[i] = g.lw[i]; out->level_height[i] = g.lh[i];
        out->level_xblocks[i] = g.lxb[i]; out->level_yblocks[i] = g.lyb[i];
    }
    for (int l = 0; l < 3; ++l) {
        out->luma_blen[l][0] = g.l[l].xblen; out->luma_blen[l][1] = g.l[l].yblen;
        out->luma_bsep[l][0] = g.l[l].xbsep; out->luma_bsep[l][1] = g.l[l].ybsep;
        out->chroma_blen[l][0] = g.c[l].xblen; out->chroma_blen[l][1] = g.c[l].yblen;
        out->chroma_bsep[l][0] = g.c[l].xbsep; out->chroma_bsep[l][1] = g.c[l].ybsep;
        out->level_factor[l] = g.level_factor[l];
    }
    return DME_OK;
}

extern "C" int oracle_downconvert(const int16_t *in, int in_w, int in_h, int16_t *out, int out_w, int out_h) {
    if (!in || !out) return fail("NULL plane");
    if (out_w < 12 || out_h < 12 || 2 * out_w > in_w || 2 * out_h > in_h) return fail("bad downconvert size");
    Plane a = make_plane(in, in_w, in_h), b(out_w, out_h);
    downconvert(a, b);
    std::memcpy(out, b.d.data(), sizeof(int16_t) * b.d.size());
    return DME_OK;
}

// CompensatePicture (mot_comp.cpp:131-186) over the three components.
static int mc_picture(const Geo &g, int prec, bool mmx, const oracle_picture *refs, int nrefs,
                      const dme_medata *me, int direction, int16_t *const planes[3], const int widths[3],
                      const int heights[3]) {
    if (nrefs < 1 || nrefs > 2 || !refs || !me || !me->vectors[0] || !me->mode || !me->sb_split)
        return fail("bad arguments");
    if (direction != DME_MC_SUBTRACT && direction != DME_MC_ADD) return fail("bad direction");
    for (int c = 0; c < 3; ++c) {
        const int pw = c ? g.CW : g.W, ph = c ? g.CH : g.H;
        if (!planes[c] || widths[c] < pw || heights[c] < ph || !me->dc[c]) return fail("bad plane");
        Plane up[2];
        const Plane *upp[2];
        for (int r = 0; r < nrefs; ++r) {
            const int16_t *src = c == 0 ? refs[r].y : (c == 1 ? refs[r].u : refs[r].v);
            if (!src) return fail("reference plane missing");
            upconvert(make_plane(src, pw, ph), up[r]);
        }
        upp[0] = &up[0];
        upp[1] = nrefs > 1 ? &up[1] : &up[0];
        McArrays m;
        m.xnb = g.xnb;
        m.ynb = g.ynb;
        m.xsb = g.xsb;
        m.nrefs = nrefs;
        m.v[0] = me->vectors[0];
        m.v[1] = nrefs > 1 ? me->vectors[1] : me->vectors[0];
        m.mode = me->mode;
        m.dc = me->dc[c];
        m.split = me->sb_split;
        mc_component(c ? g.c[2] : g.l[2], prec, mmx, pw, ph, c ? 1 : 0, m, upp, direction, planes[c], widths[c],
                     heights[c]);
    }
    return 0;
}

extern "C" int oracle_motion_compensate(const dme_params *p, int width, int height, const oracle_picture *refs,
                                        int nrefs, const dme_medata *me, int direction, int16_t *const planes[3],
                                        const int widths[3], const int heights[3]) {
    if (p && p->mv_precision == 3) {  // eighth pel: compensation only, geometry as quarter pel
        dme_params q = *p;
        q.mv_precision = 2;
        Geo g;
        if (make_geo(&q, width, height, g)) return -1;
        return mc_picture(g, 3, q.fp_semantics == DME_FP_MMX, refs, nrefs, me, direction, planes, widths, heights);
    }
    Geo g;
    if (make_geo(p, width, height, g)) return -1;
    return mc_picture(g, p->mv_precision, p->fp_semantics == DME_FP_MMX, refs, nrefs, me, direction, planes, widths, heights);
}

extern "C" int oracle_combine(const int16_t *y, const int16_t *u, const int16_t *v, int w, int h, int16_t *out) {
    if (!y || !u || !v || !out || w < 2 || h < 2 || (w & 1) || (h & 1)) return fail("bad arguments");
    const Plane c = combine(y, u, v, w, h);
    std::memcpy(out, c.d.data(), sizeof(int16_t) * c.d.size());
    return DME_OK;
}

extern "C" int oracle_upconvert(const int16_t *in, int w, int h, int16_t *out) {
    if (!in || !out) return fail("NULL plane");
    if (w < 8 || h < 8) return fail("upconvert needs at least 8x8");
    Plane a = make_plane(in, w, h), u;
    upconvert(a, u);
    std::memcpy(out, u.d.data(), sizeof(int16_t) * u.d.size());
    return DME_OK;
}

extern "C" int oracle_lambda_map(int xnb, int ynb, int num_refs, const dme_mv *v1, const dme_mv *v2, float lambda,
                                 float *out) {
    if (xnb < 3 || ynb < 3 || !v1 || !out || (num_refs > 1 && !v2)) return fail("bad lambda-map arguments");
    std::vector<MV> a((size_t)xnb * ynb), b;
    for (size_t k = 0; k < a.size(); ++k) a[k] = MV(v1[k].x, v1[k].y);
    if (num_refs > 1) {
        b.resize(a.size());
        for (size_t k = 0; k < b.size(); ++k) b[k] = MV(v2[k].x, v2[k].y);
    }
    lambda_map(xnb, ynb, num_refs, a.data(), num_refs > 1 ? b.data() : nullptr, lambda, out);
    return DME_OK;
}

extern "C" int oracle_me(const dme_params *p, const oracle_picture *cur, const oracle_picture *refs, int nrefs,
                         int stop_after, const float *lambda_override, dme_medata *out, oracle_capture *cap,
                         double *sad_terms) {
    if (!cur || !cur->y || !refs || nrefs < 1 || nrefs > 2) return fail("bad picture arguments");
    Geo g;
    int rc = make_geo(p, cur->width, cur->height, g);
    if (rc) return rc;
    for (int r = 0; r < nrefs; ++r)
        if (!refs[r].y || refs[r].width != cur->width || refs[r].height != cur->height)
            return fail("reference size differs from the picture");
    if (nrefs == 2 && refs[0].pic_num == refs[1].pic_num) return fail("two references with the same picture number");
    double terms = 0.0;
    const int W = g.W, H = g.H;

    // the picture ME works on: luma, or the combined-component picture
    // (EncPicture::DataForME, enc_picture.cpp:114-120); SetDC reads luma
    const Plane luma = make_plane(cur->y, W, H);
    Plane pic = luma;
    Plane ref[2];
    for (int r = 0; r < nrefs; ++r) ref[r] = make_plane(refs[r].y, W, H);
    if (p->combined_me) {
        if (!cur->u || !cur->v) return fail("combined-component ME needs the picture's chroma");
        for (int r = 0; r < nrefs; ++r)
            if (!refs[r].u || !refs[r].v) return fail("combined-component ME needs the references' chroma");
        pic = combine(cur->y, cur->u, cur->v, W, H);
        for (int r = 0; r < nrefs; ++r) ref[r] = combine(refs[r].y, refs[r].u, refs[r].v, W, H);
    }

    MEData me;
    me.init(g.xnb, g.ynb, g.xsb, g.ysb, nrefs);

    // ---- PixelME: PixelMatcher::DoSearch (pixel_match.cpp:59-157) ----
    {
        const int D = g.depth;
        std::vector<Plane> pd(D + 1), rd[2];
        pd[0] = pic;
        for (int i = 1; i <= D; ++i) { pd[i] = Plane(g.lw[i], g.lh[i]); downconvert(pd[i - 1], pd[i]); }
        for (int r = 0; r < nrefs; ++r) {
            rd[r].resize(D + 1);
            rd[r][0] = ref[r];
            for (int i = 1; i <= D; ++i) { rd[r][i] = Plane(g.lw[i], g.lh[i]); downconvert(rd[r][i - 1], rd[r][i]); }
        }
        if (cap)
            for (int i = 1; i <= D; ++i)
                if (cap->cur_pyramid[i]) std::memcpy(cap->cur_pyramid[i], pd[i].d.data(), sizeof(int16_t) * pd[i].d.size());
        for (int r = 0; r < nrefs; ++r) {
            std::vector<PelLevel> lv(D + 1);
            for (int l = D; l >= 0; --l) {
                lv[l].xnb = g.lxb[l];
                lv[l].ynb = g.lyb[l];
                int fxr = -1, fyr = -1;
                if (p->full_search) {
                    // m_tdiff, m_big_xr (pixel_match.cpp:79-80, 249-250)
                    const int td = std::abs(refs[r].pic_num - cur->pic_num);
                    fxr = std::min(td, 3) * p->x_range_me;
                    fyr = std::min(td, 3) * p->y_range_me;
                }
                match_pic_pel(pd[l], rd[r][l], g.l[2], l, l < D ? &lv[l + 1] : nullptr, lv[l], terms, fxr, fyr);
                if (cap && cap->level_vectors[r][l])
                    for (size_t k = 0; k < lv[l].vec.size(); ++k) {
                        cap->level_vectors[r][l][k].x = lv[l].vec[k].x;
                        cap->level_vectors[r][l][k].y = lv[l].vec[k].y;
                    }
                if (cap && cap->level_costs[r][l])
                    for (size_t k = 0; k < lv[l].pc.size(); ++k) {
                        cap->level_costs[r][l][k].SAD = lv[l].pc[k].SAD;
                        cap->level_costs[r][l][k].mvcost = lv[l].pc[k].mvcost;
                        cap->level_costs[r][l][k].total = lv[l].pc[k].total;
                    }
            }
            me.vec[r] = lv[0].vec;
            me.pc[r] = lv[0].pc;
        }
    }
    if (stop_after <= ORACLE_STOP_PEL) {
        export_medata(me, out);
        if (sad_terms) *sad_terms = terms;
        return DME_OK;
    }

    // ---- SubPixelME: lambda map (picture_compress.cpp:208-226) ----
    if (lambda_override) {
        std::memcpy(me.lam.data(), lambda_override, sizeof(float) * me.lam.size());
    } else {
        lambda_map(g.xnb, g.ynb, nrefs, me.vec[0].data(), nrefs > 1 ? me.vec[1].data() : nullptr, p->lambda,
                   me.lam.data());
    }
    if (stop_after <= ORACLE_STOP_LAMBDA) {
        export_medata(me, out);
        if (sad_terms) *sad_terms = terms;
        return DME_OK;
    }

    // ---- SubPixelME: refinement or pixel-precision doubling
    //      (picture_compress.cpp:228-259) ----
    Plane up[2];
    for (int r = 0; r < nrefs; ++r) {
        upconvert(ref[r], up[r]);
        if (cap && cap->ref_up[r]) std::memcpy(cap->ref_up[r], up[r].d.data(), sizeof(int16_t) * up[r].d.size());
    }
    int prec = p->mv_precision;
    if (prec != 0) {
        for (int r = 0; r < nrefs; ++r) subpel_ref(pic, up[r], g, prec, p->fp_semantics, me, r, terms);
    } else {
        for (int r = 0; r < nrefs; ++r)
            for (auto &v : me.vec[r]) v = MV(v.x << 1, v.y << 1);
        prec = 1;  // SetMVPrecision(MV_PRECISION_HALF_PIXEL)
    }
    if (stop_after <= ORACLE_STOP_SUBPEL) {
        // state as the reference holds it between SubPixelME and
        // ModeDecisionME (pixel-precision vectors still doubled)
        export_medata(me, out);
        if (sad_terms) *sad_terms = terms;
        return DME_OK;
    }

    // ---- ModeDecisionME: ModeDecider::DoModeDecn (me_mode_decn.cpp:66-177) ----
    {
        MEData l0, l1;
        l0.init(g.xnb / 4, g.ynb / 4, g.xsb, g.ysb, nrefs);
        l1.init(g.xnb / 2, g.ynb / 2, g.xsb, g.ysb, nrefs);
        level_lambda(me, 0, (float)(1.0 / g.level_factor[0]), l0.xnb, l0.ynb, l0.lam);
        level_lambda(me, 1, (float)(1.0 / g.level_factor[1]), l1.xnb, l1.ynb, l1.lam);
        ModeCtx m;
        m.g = &g; m.pic = &pic; m.up[0] = &up[0]; m.up[1] = nrefs > 1 ? &up[1] : &up[0];
        m.nrefs = nrefs; m.step = prec; m.fp_sem = p->fp_semantics;
        m.lv[0] = &l0; m.lv[1] = &l1; m.lv[2] = &me; m.terms = &terms;
        for (int ys = 0; ys < g.ysb; ++ys)
            for (int xs = 0; xs < g.xsb; ++xs) {
                // DoSBDecn (me_mode_decn.cpp:179-197): level 0 always runs,
                // because SBCosts after level 1 is never above the level-2 cost.
                mode_level(m, xs, ys, 2);
                const float old_best = me.sbc[(size_t)ys * g.xsb + xs];
                mode_level(m, xs, ys, 1);
                if (me.sbc[(size_t)ys * g.xsb + xs] <= old_best) mode_level(m, xs, ys, 0);
            }
        // SetDC (me_mode_decn.cpp:567-573) on the original Y/U/V
        set_dc(luma, g.l[2], me, 0);
        if (cur->u && cur->v) {
            Plane pu = make_plane(cur->u, g.CW, g.CH), pv = make_plane(cur->v, g.CW, g.CH);
            set_dc(pu, g.c[2], me, 1);
            set_dc(pv, g.c[2], me, 2);
        }
        terms += 3.0 * W * H / 2.0;  // DC of every Y, U and V sample
    }
    if (p->mv_precision == 0)
        for (int r = 0; r < nrefs; ++r)
            for (auto &v : me.vec[r]) v = MV(v.x >> 1, v.y >> 1);

    export_medata(me, out);
    if (sad_terms) *sad_terms = terms;
    return DME_OK;
}

// The fast form of the learned path from level-0 integer vectors v0[r]
// (B each; blocks with no samples must hold (0, 0)): pel SADs, lambda map
// (MEData::SetLambdaMap), the independent sub-pel refinement and mode
// decision, the exact-cost pass and SetDC.  v2_out[r] (may be NULL) receive
// the pre-merge sub-pel vectors.  Quarter- or half-pel precision, MMX
// float semantics.
extern "C" int oracle_fast_me(const dme_params *p, const oracle_picture *cur, const oracle_picture *refs, int nrefs,
                              const dme_mv *const *v0, dme_medata *out, dme_mv *const *v2_out, double *sad_terms) {
    if (!cur || !cur->y || !refs || !v0 || nrefs < 1 || nrefs > 2) return fail("bad arguments");
    if (p->mv_precision < 1 || p->mv_precision > 2 || p->fp_semantics != DME_FP_MMX || p->full_search ||
        p->combined_me)
        return fail("the fast form runs half or quarter pel, MMX semantics, hierarchical search on luma");
    Geo g;
    int rc = make_geo(p, cur->width, cur->height, g);
    if (rc) return rc;
    const int W = g.W, H = g.H, prec = p->mv_precision;
    double terms = 0.0;
    Plane pic = make_plane(cur->y, W, H);
    Plane ref[2], up[2];
    for (int r = 0; r < nrefs; ++r) {
        if (!refs[r].y || refs[r].width != W || refs[r].height != H) return fail("reference size differs");
        ref[r] = make_plane(refs[r].y, W, H);
        upconvert(ref[r], up[r]);
    }
    MEData me;
    me.init(g.xnb, g.ynb, g.xsb, g.ysb, nrefs);
    const size_t B = (size_t)g.xnb * g.ynb;
    for (int r = 0; r < nrefs; ++r)
        for (int y = 0; y < g.ynb; ++y)
            for (int x = 0; x < g.xnb; ++x) {
                const size_t k = me.idx(y, x);
                const Blk b = block_limits(g.l[2], W, H, x, y);
                me.vec[r][k] = MV(v0[r][k].x, v0[r][k].y);
                me.pc[r][k].SAD = b.empty() ? 10000000.0f : (float)pel_sad(pic, ref[r], b, me.vec[r][k]);
                if (!b.empty()) terms += (double)b.xl * b.yl;
            }
    lambda_map(g.xnb, g.ynb, nrefs, me.vec[0].data(), nrefs > 1 ? me.vec[1].data() : nullptr, p->lambda,
               me.lam.data());
    std::vector<MV> V0[2];
    std::vector<float> PS[2];
    for (int r = 0; r < nrefs; ++r) {
        V0[r] = me.vec[r];
        PS[r].resize(B);
        for (size_t k = 0; k < B; ++k) PS[r][k] = me.pc[r][k].SAD;
    }
    // independent sub-pel refinement: predictors from the integer vectors
    for (int r = 0; r < nrefs; ++r) {
        std::vector<MV> pf(B);
        for (size_t k = 0; k < B; ++k) pf[k] = MV(me.vec[r][k].x << prec, me.vec[r][k].y << prec);
        subpel_ref(pic, up[r], g, prec, FP_ONE_ROUNDING, me, r, terms, &pf);
    }
    std::vector<MV> V2[2];
    for (int r = 0; r < nrefs; ++r) {
        V2[r] = me.vec[r];
        if (v2_out && v2_out[r])
            for (size_t k = 0; k < B; ++k) { v2_out[r][k].x = V2[r][k].x; v2_out[r][k].y = V2[r][k].y; }
        fast_l2_costs(pic, up[r], g, prec, me, r, V0[r], PS[r], terms);
    }
    MEData l0, l1;
    l0.init(g.xnb / 4, g.ynb / 4, g.xsb, g.ysb, nrefs);
    l1.init(g.xnb / 2, g.ynb / 2, g.xsb, g.ysb, nrefs);
    level_lambda(me, 0, (float)(1.0 / g.level_factor[0]), l0.xnb, l0.ynb, l0.lam);
    level_lambda(me, 1, (float)(1.0 / g.level_factor[1]), l1.xnb, l1.ynb, l1.lam);
    ModeCtx m;
    m.g = &g; m.pic = &pic; m.up[0] = &up[0]; m.up[1] = nrefs > 1 ? &up[1] : &up[0];
    m.nrefs = nrefs; m.step = prec; m.fp_sem = FP_ONE_ROUNDING;
    m.lv[0] = &l0; m.lv[1] = &l1; m.lv[2] = &me; m.terms = &terms;
    m.pred_v2[0] = &V2[0];
    m.pred_v2[1] = nrefs > 1 ? &V2[1] : nullptr;
    for (int ys = 0; ys < g.ysb; ++ys)
        for (int xs = 0; xs < g.xsb; ++xs) {
            mode_level(m, xs, ys, 2);
            mode_level(m, xs, ys, 1);
            mode_level(m, xs, ys, 0);
        }
    for (int ys = 0; ys < g.ysb; ++ys)
        for (int xs = 0; xs < g.xsb; ++xs) fast_eval_sb(m, V2, xs, ys);
    set_dc(pic, g.l[2], me, 0);
    if (cur->u && cur->v) {
        Plane pu = make_plane(cur->u, g.CW, g.CH), pv = make_plane(cur->v, g.CW, g.CH);
        set_dc(pu, g.c[2], me, 1);
        set_dc(pv, g.c[2], me, 2);
    }
    export_medata(me, out);
    if (sad_terms) *sad_terms = terms;
    return DME_OK;
}

// The exact-cost pass alone, on any final decisions: v2[r] the pre-merge
// sub-pel vectors, final_vec[r] / split / lambda_map the final MEData.
// Writes the level-2 costs (pc[r], B each) and the superblock costs (sbc, S)
// that the reference's Lagrangian gives those decisions; for the plain
// search's own decisions they are its PredCosts and SBCosts.
extern "C" int oracle_exact_costs(const dme_params *p, const oracle_picture *cur, const oracle_picture *refs,
                                  int nrefs, const dme_mv *const *v0, const dme_mv *const *v2,
                                  const dme_mv *const *final_vec,
                                  const int32_t *split, const float *lambda_map, dme_cost *const *pc, float *sbc,
                                  int32_t *mode) {
    Geo g;
    int rc = make_geo(p, cur->width, cur->height, g);
    if (rc) return rc;
    const int W = g.W, H = g.H, prec = p->mv_precision == 0 ? 1 : p->mv_precision;
    double terms = 0.0;
    Plane pic = make_plane(cur->y, W, H);
    Plane up[2];
    for (int r = 0; r < nrefs; ++r) upconvert(make_plane(refs[r].y, W, H), up[r]);
    MEData me;
    me.init(g.xnb, g.ynb, g.xsb, g.ysb, nrefs);
    const size_t B = (size_t)g.xnb * g.ynb, S = (size_t)g.xsb * g.ysb;
    std::memcpy(me.lam.data(), lambda_map, sizeof(float) * B);
    std::vector<MV> V2[2], V0[2];
    std::vector<float> PS[2];
    for (int r = 0; r < nrefs; ++r) {
        const Plane rf = make_plane(refs[r].y, W, H);
        me.vec[r].resize(B);
        V2[r].resize(B);
        V0[r].resize(B);
        PS[r].resize(B);
        for (size_t k = 0; k < B; ++k) {
            V2[r][k] = MV(v2[r][k].x, v2[r][k].y);
            V0[r][k] = MV(v0[r][k].x, v0[r][k].y);
            me.vec[r][k] = V2[r][k];
            const Blk b = block_limits(g.l[2], W, H, (int)(k % g.xnb), (int)(k / g.xnb));
            PS[r][k] = b.empty() ? 10000000.0f : (float)pel_sad(pic, rf, b, V0[r][k]);
        }
        fast_l2_costs(pic, up[r], g, prec, me, r, V0[r], PS[r], terms);
        for (size_t k = 0; k < B; ++k) {
            pc[r][k].SAD = me.pc[r][k].SAD;
            pc[r][k].mvcost = me.pc[r][k].mvcost;
            pc[r][k].total = me.pc[r][k].total;
            me.vec[r][k] = MV(final_vec[r][k].x, final_vec[r][k].y);
        }
    }
    MEData l0, l1;
    l0.init(g.xnb / 4, g.ynb / 4, g.xsb, g.ysb, nrefs);
    l1.init(g.xnb / 2, g.ynb / 2, g.xsb, g.ysb, nrefs);
    level_lambda(me, 0, (float)(1.0 / g.level_factor[0]), l0.xnb, l0.ynb, l0.lam);
    level_lambda(me, 1, (float)(1.0 / g.level_factor[1]), l1.xnb, l1.ynb, l1.lam);
    ModeCtx m;
    m.g = &g; m.pic = &pic; m.up[0] = &up[0]; m.up[1] = nrefs > 1 ? &up[1] : &up[0];
    m.nrefs = nrefs; m.step = prec; m.fp_sem = DME_FP_MMX;
    m.lv[0] = &l0; m.lv[1] = &l1; m.lv[2] = &me; m.terms = &terms;
    // level-2 superblock costs: DoLevelDecn(2) on the exact block costs
    for (int ys = 0; ys < g.ysb; ++ys)
        for (int xs = 0; xs < g.xsb; ++xs) {
            const size_t si = (size_t)ys * g.xsb + xs;
            std::vector<MV> keep[2];
            // DoUnitDecn(2) reads the level-2 vectors for the bi-prediction: the pre-merge ones
            for (int r = 0; r < nrefs; ++r)
                for (int v = 4 * ys; v < 4 * ys + 4; ++v)
                    for (int u = 4 * xs; u < 4 * xs + 4; ++u) {
                        keep[r].push_back(me.vec[r][me.idx(v, u)]);
                        me.vec[r][me.idx(v, u)] = V2[r][me.idx(v, u)];
                    }
            mode_level(m, xs, ys, 2);
            for (int r = 0; r < nrefs; ++r) {
                size_t q = 0;
                for (int v = 4 * ys; v < 4 * ys + 4; ++v)
                    for (int u = 4 * xs; u < 4 * xs + 4; ++u) me.vec[r][me.idx(v, u)] = keep[r][q++];
            }
            me.split[si] = split[si];
        }
    for (int ys = 0; ys < g.ysb; ++ys)
        for (int xs = 0; xs < g.xsb; ++xs) fast_eval_sb(m, V2, xs, ys);
    std::memcpy(sbc, me.sbc.data(), sizeof(float) * S);
    if (mode) for (size_t k = 0; k < B; ++k) mode[k] = me.mode[k];
    return DME_OK;
}
