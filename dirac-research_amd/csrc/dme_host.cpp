// dme_host.cpp -- host-side parts of the engine that need no GPU: error
// reporting, parameter rationalisation / geometry, and the lambda map.
//
// The lambda map stays on the host on purpose: the reference derives its
// thresholds from x87 80-bit long double statistics (motion.cpp:201-333),
// which the GPU has no type for; the map is tiny (one float per block).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "dme_internal.h"

namespace dme {

static thread_local std::string t_last_error;

int set_error(int code, const std::string &msg) {
    t_last_error = msg;
    return code;
}

int exception_status() noexcept {
    auto note = [](const char *a, const char *b) {
        try {
            t_last_error = a;
            t_last_error += b;
        } catch (...) {  // no memory for the message: keep the status
        }
    };
    try {
        throw;
    } catch (const std::bad_alloc &) {
        note("host allocation failed", "");
        return DME_ERR_NOMEM;
    } catch (const std::exception &e) {
        note("unexpected error: ", e.what());
        return DME_ERR_INVALID;
    } catch (...) {
        note("unexpected error", "");
        return DME_ERR_INVALID;
    }
}

// PicturePredParams::SetBlockSizes (common.cpp:207-320) for 4:2:0, then
// SetMotionParameters (seq_compress.cpp:124-132), the pyramid depth of
// PixelMatcher::DoSearch (pixel_match.cpp:94-95) and the MEData grids of
// MakeMEDataHierarchy (pixel_match.cpp:185-220).
int make_geometry(const dme_params &p, int width, int height, Geometry &g) {
    if (p.chroma_format != 0) return set_error(DME_ERR_INVALID, "only 4:2:0 chroma is supported");
    if (p.mv_precision < 0 || p.mv_precision > 3)
        return set_error(DME_ERR_INVALID, "mv_precision must be 0 (pel), 1 (half), 2 (quarter) or 3 (eighth)");
    if (p.fp_semantics != DME_FP_MMX && p.fp_semantics != DME_FP_C)
        return set_error(DME_ERR_INVALID, "fp_semantics must be DME_FP_MMX or DME_FP_C");
    if (p.combined_me != 0 && p.combined_me != 1) return set_error(DME_ERR_INVALID, "combined_me must be 0 or 1");
    for (int i = 0; i < 3; ++i)
        if (p.reserved[i]) return set_error(DME_ERR_INVALID, "reserved params must be zero");
    if (p.full_search != 0 && p.full_search != 1) return set_error(DME_ERR_INVALID, "full_search must be 0 or 1");
    if (p.full_search && (p.x_range_me < 0 || p.y_range_me < 0 || p.x_range_me > 128 || p.y_range_me > 128))
        return set_error(DME_ERR_INVALID, "full-search ranges must be in [0, 128]");
    if (!p.full_search && (p.x_range_me || p.y_range_me))
        return set_error(DME_ERR_INVALID, "x_range_me / y_range_me apply to full search only");
    if (p.xblen <= 0 || p.yblen <= 0 || p.xbsep <= 0 || p.ybsep <= 0)
        return set_error(DME_ERR_INVALID, "block lengths and separations must be positive");
    if (width < 24 || height < 24 || width > 16384 || height > 16384 || (width & 1) || (height & 1))
        return set_error(DME_ERR_INVALID, "picture size must be even and within [24, 16384]");
    if (!(p.lambda >= 0.0f) || p.lambda > 1e6f) return set_error(DME_ERR_INVALID, "lambda out of range");

    int xblen = p.xblen, yblen = p.yblen, xbsep = p.xbsep, ybsep = p.ybsep;
    int rem = xbsep % 4;
    if (rem != 0) { xbsep += 4 - rem; xblen = xbsep + 4; }
    rem = ybsep % 4;
    if (rem != 0) { ybsep += 4 - rem; yblen = ybsep + 4; }
    if (xblen % 4 != 0) xblen = xbsep + 4;
    if (yblen % 4 != 0) yblen = ybsep + 4;
    if (xbsep > xblen) xblen = xbsep + 4;
    if (ybsep > yblen) yblen = ybsep + 4;
    if (2 * xbsep < xblen) xblen = xbsep + 4;
    if (2 * ybsep < yblen) yblen = ybsep + 4;
    if (xblen > 32 || yblen > 32) return set_error(DME_ERR_INVALID, "block lengths above 32 are not supported");

    g.W = width; g.H = height; g.CW = width >> 1; g.CH = height >> 1;
    g.blen[2][0] = xblen; g.blen[2][1] = yblen; g.bsep[2][0] = xbsep; g.bsep[2][1] = ybsep;
    // sub-SB and SB: both y lengths add the x separation (common.cpp:297,302)
    for (int k = 1; k >= 0; --k) {
        g.bsep[k][0] = g.bsep[k + 1][0] * 2;
        g.blen[k][0] = g.blen[k + 1][0] + g.bsep[k + 1][0];
        g.bsep[k][1] = g.bsep[k + 1][1] * 2;
        g.blen[k][1] = g.blen[k + 1][1] + g.bsep[k + 1][0];
    }
    for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 2; ++a) g.boff[k][a] = (g.blen[k][a] - g.bsep[k][a]) / 2;
    g.cblen[0] = xblen / 2; g.cblen[1] = yblen / 2; g.cbsep[0] = xbsep / 2; g.cbsep[1] = ybsep / 2;
    for (int a = 0; a < 2; ++a) g.cboff[a] = (g.cblen[a] - g.cbsep[a]) / 2;

    g.xsb = (width + g.bsep[0][0] - 1) / g.bsep[0][0];
    g.ysb = (height + g.bsep[0][1] - 1) / g.bsep[0][1];
    g.xnb = 4 * g.xsb;
    g.ynb = 4 * g.ysb;

    if (p.full_search)
        g.depth = 0;  // no hierarchy (pixel_match.cpp:91, 140-155)
    else if (p.depth < 0)
        g.depth = (int)std::min(std::log(((double)width) / 12.0) / std::log(2.0),
                                std::log(((double)height) / 12.0) / std::log(2.0));
    else
        g.depth = p.depth;
    if (!p.full_search && (g.depth < 1 || g.depth > DME_MAX_DEPTH))
        return set_error(DME_ERR_INVALID, "pyramid depth must be in [1, 8]");
    g.lw[0] = width; g.lh[0] = height; g.lxb[0] = g.xnb; g.lyb[0] = g.ynb;
    int sf = 1, xn = g.xnb, yn = g.ynb;
    for (int i = 1; i <= g.depth; ++i) {
        sf *= 2;
        g.lw[i] = width / sf;
        g.lh[i] = height / sf;
        xn >>= 1; yn >>= 1;
        if (g.lw[i] % xbsep != 0) xn++;
        if (g.lh[i] % ybsep != 0) yn++;
        g.lxb[i] = xn; g.lyb[i] = yn;
        if (g.lw[i] < 12 || g.lh[i] < 12)
            return set_error(DME_ERR_INVALID, "pyramid too deep: a level is smaller than 12 samples");
        if (xn < 2) return set_error(DME_ERR_INVALID, "pyramid too deep: a level has fewer than 2 block columns");
    }
    g.level_factor[0] = float(16 * xblen * yblen) / float(g.blen[0][0] * g.blen[0][1]);
    g.level_factor[1] = float(4 * xblen * yblen) / float(g.blen[1][0] * g.blen[1][1]);
    g.level_factor[2] = 1.0f;
    return DME_OK;
}

// MEData::FindTransitions, vector-variation part (motion.cpp:258-331); the
// SAD-based map computed before it is overwritten in full and has no effect.
static void transitions(int X, int Y, const dme_mv *mv, std::vector<unsigned char> &t) {
    std::vector<double> val((size_t)X * Y, 0.0);
    long double total = 0.0;
    for (int j = 1; j < Y - 1; ++j)
        for (int i = 1; i < X - 1; ++i) {
            const dme_mv c = mv[(size_t)j * X + i];
            double v = 0.0;
            for (int q = -1; q <= 1; ++q)
                for (int p = -1; p <= 1; ++p) {
                    const dme_mv n = mv[(size_t)(j + q) * X + (i + p)];
                    v = std::max(v, (double)(std::abs(n.x - c.x) + std::abs(n.y - c.y)));
                }
            val[(size_t)j * X + i] = v;
            total += v;
        }
    const long double mean = total / static_cast<long double>(X * Y);
    long double sd = 0.0;
    for (int j = 1; j < Y - 1; ++j)
        for (int i = 1; i < X - 1; ++i) {
            double d = val[(size_t)j * X + i] - mean;
            d *= d;
            sd += d;
        }
    sd /= static_cast<long double>(X * Y);
    sd = std::sqrt(sd);
    const double thr = mean + 3 * sd;
    t.resize((size_t)X * Y);
    for (size_t k = 0; k < t.size(); ++k) t[k] = val[k] >= thr;
    for (int j = 0; j < Y / 4; ++j)
        for (int i = 0; i < X / 4; ++i) {
            unsigned char any = 0;
            for (int q = 4 * j; q < 4 * j + 4; ++q)
                for (int p = 4 * i; p < 4 * i + 4; ++p) any |= t[(size_t)q * X + p];
            for (int q = 4 * j; q < 4 * j + 4; ++q)
                for (int p = 4 * i; p < 4 * i + 4; ++p) t[(size_t)q * X + p] = any;
        }
}

// MEData::SetLambdaMap(num_refs, lambda) (motion.cpp:128-171)
void lambda_map_l2(int X, int Y, int nrefs, const dme_mv *v1, const dme_mv *v2, float lambda, float *out) {
    std::vector<unsigned char> t1, t2;
    transitions(X, Y, v1, t1);
    if (nrefs > 1) transitions(X, Y, v2, t2);
    for (int j = 0; j < Y; ++j)
        for (int i = 0; i < X; ++i) {
            const size_t k = (size_t)j * X + i;
            float m;
            if (nrefs == 1) m = t1[k] ? 0.0f : lambda;
            else if (t1[k] && t2[k]) m = 0.0f;
            else if (t1[k] || t2[k]) m = (float)(lambda / 4.0);
            else m = lambda;
            if (i < 4 || j < 4) m = (float)(m / 5.0);
            out[k] = m;
        }
}

// MEData::SetLambdaMap(level, map, wt) (motion.cpp:173-199)
void lambda_map_level(const float *l2, int X2, int level, float wt, int Xl, int Yl, float *out) {
    const int f = 1 << (2 - level);
    for (int j = 0; j < Yl; ++j)
        for (int i = 0; i < Xl; ++i) {
            float m = l2[(size_t)(f * j) * X2 + f * i];
            for (int q = f * j; q < f * (j + 1); ++q)
                for (int p = f * i; p < f * (i + 1); ++p) m = std::max(l2[(size_t)q * X2 + p], m);
            out[(size_t)j * Xl + i] = m * wt;
        }
}

}  // namespace dme

extern "C" const char *dme_last_error(void) { return dme::t_last_error.c_str(); }

extern "C" int dme_version(void) { return 100; }

extern "C" int dme_geometry_for(const dme_params *params, int width, int height, dme_geometry *out) try {
    if (!params || !out) return dme::set_error(DME_ERR_INVALID, "NULL argument");
    dme::Geometry g;
    int rc = dme::make_geometry(*params, width, height, g);
    if (rc) return rc;
    std::memset(out, 0, sizeof(*out));
    out->width = g.W; out->height = g.H; out->chroma_width = g.CW; out->chroma_height = g.CH;
    out->xnum_sb = g.xsb; out->ynum_sb = g.ysb; out->xnum_blocks = g.xnb; out->ynum_blocks = g.ynb;
    out->depth = g.depth;
    for (int i = 0; i <= g.depth; ++i) {
        out->level_width[i] = g.lw[i]; out->level_height[i] = g.lh[i];
        out->level_xblocks[i] = g.lxb[i]; out->level_yblocks[i] = g.lyb[i];
    }
    for (int l = 0; l < 3; ++l) {
        for (int a = 0; a < 2; ++a) {
            out->luma_blen[l][a] = g.blen[l][a];
            out->luma_bsep[l][a] = g.bsep[l][a];
        }
        out->level_factor[l] = g.level_factor[l];
    }
    // chroma parameters of all levels (common.cpp:278-313)
    int cbl[3][2], cbs[3][2];
    cbl[2][0] = g.cblen[0]; cbl[2][1] = g.cblen[1]; cbs[2][0] = g.cbsep[0]; cbs[2][1] = g.cbsep[1];
    for (int k = 1; k >= 0; --k) {
        cbs[k][0] = cbs[k + 1][0] * 2; cbl[k][0] = cbl[k + 1][0] + cbs[k + 1][0];
        cbs[k][1] = cbs[k + 1][1] * 2; cbl[k][1] = cbl[k + 1][1] + cbs[k + 1][0];
    }
    for (int l = 0; l < 3; ++l)
        for (int a = 0; a < 2; ++a) { out->chroma_blen[l][a] = cbl[l][a]; out->chroma_bsep[l][a] = cbs[l][a]; }
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_lambda_map_host(int xnum_blocks, int ynum_blocks, int num_refs, const dme_mv *vectors1,
                                   const dme_mv *vectors2, float lambda, float *out_lambda_map) try {
    if (xnum_blocks < 3 || ynum_blocks < 3 || !vectors1 || !out_lambda_map || num_refs < 1 || num_refs > 2 ||
        (num_refs == 2 && !vectors2))
        return dme::set_error(DME_ERR_INVALID, "bad lambda-map arguments");
    dme::lambda_map_l2(xnum_blocks, ynum_blocks, num_refs, vectors1, vectors2, lambda, out_lambda_map);
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}
