// dme_mvcodec.cpp -- the MV-data wire format of an inter picture: the bytes
// PictureCompressor::CodeMVData writes (libdirac_encoder/picture_compress.cpp:
// 463-530) and PictureDecompressor::DecompressMVData reads (libdirac_decoder/
// picture_decompress.cpp:266-357).  Host code, no GPU: the coder is a chain of
// adaptive binary decisions per picture.
//
// Layout (libdirac_byteio/mvdata_byteio.cpp, mvdataelement_byteio.cpp):
//   header: block-parameter index (and the four lengths when 0), MV
//     precision, a zero bit (no global motion), 0 (frame prediction mode),
//     the picture-weights flag (and precision and weights when set), byte
//     alignment; every number an interleaved exp-Golomb code (ByteIO::
//     WriteUint / WriteSint, byteio.cpp);
//   then eight (two references) or six elements in the order split modes,
//     prediction modes, reference 1 x and y, [reference 2 x and y], DC of
//     Y, U, V: each the byte count of its data as an exp-Golomb code, byte
//     alignment, then its arithmetic-coded data (mv_codec.cpp:49-641 over
//     the coder of arith_codec.h/.cpp).
// Elements are coded per superblock in raster order; inside a superblock,
// per prediction unit (1, 4 or 16 blocks by the split) in raster order, the
// unit's top-left block carrying its values.
#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "dme_internal.h"

namespace {

using dme::set_error;

#define MV_TRY(expr)                   \
    do {                               \
        const int rc_ = (expr);        \
        if (rc_ != DME_OK) return rc_; \
    } while (0)

// ---- bits ---------------------------------------------------------------------

// Bits written most significant first into bytes (ByteIO::WriteBit); align()
// emits a partial byte padded with zeros (ByteIO::ByteAlignOutput).
struct BitSink {
    std::vector<uint8_t> bytes;
    unsigned cur = 0;
    int used = 0;
    void put(bool b) {
        cur |= (unsigned)b << (7 - used);
        if (++used == 8) {
            bytes.push_back((uint8_t)cur);
            cur = 0;
            used = 0;
        }
    }
    void align() {
        if (used) {
            bytes.push_back((uint8_t)cur);
            cur = 0;
            used = 0;
        }
    }
    // interleaved exp-Golomb: for value + 1 = 1 b_(n-1) .. b_0, the pairs
    // (0, b_i) from the top, then 1
    void put_uint(unsigned v) {
        const unsigned long long w = (unsigned long long)v + 1;
        int n = 0;
        while ((w >> (n + 1)) != 0) ++n;
        for (int i = n - 1; i >= 0; --i) {
            put(false);
            put((w >> i) & 1);
        }
        put(true);
    }
    void put_sint(int v) {
        put_uint(v < 0 ? 0u - (unsigned)v : (unsigned)v);
        if (v) put(v < 0);
    }
};

// Reader of the same codes (ByteIO::ReadBit/ReadUint/ReadSint); reading
// past the end is an error.
struct BitSource {
    const uint8_t *p;
    size_t n, pos = 0;  // pos in bits
    bool fail = false;
    BitSource(const uint8_t *q, size_t len) : p(q), n(len) {}
    bool get() {
        if (pos >= 8 * n) {
            fail = true;
            return false;
        }
        const bool b = (p[pos >> 3] >> (7 - (pos & 7))) & 1;
        ++pos;
        return b;
    }
    void align() { pos = (pos + 7) & ~(size_t)7; }
    unsigned get_uint() {
        unsigned long long v = 1;
        while (!get()) {
            if (fail || v > 0xffffffffull) {
                fail = true;
                return 0;
            }
            v = (v << 1) | (unsigned)get();
        }
        if (v - 1 > 0xffffffffull) fail = true;
        return (unsigned)(v - 1);
    }
    int get_sint() {
        const unsigned m = get_uint();
        if (m > 0x7fffffffu) {
            fail = true;
            return 0;
        }
        return m && get() ? -(int)m : (int)m;
    }
};

// ---- the adaptive binary arithmetic coder ----------------------------------------

// Context indices of the MV data (enum MvCtxAliases, common.h:130-169).
enum {
    kDcBin1 = 0, kDcBin2 = 1,                 // info 2, sign 3
    kMvBin1 = 4, kMvBin5 = 8,                 // bins 4..8, info 9, sign 10
    kPmodeBit0 = 11, kPmodeBit1 = 12,
    kSplitBin1 = 13, kSplitBin2 = 14,         // info 15
    kContexts = 16
};

// Probability-of-zero adaptation table (Context::lut, arith_codec.cpp:42-75):
// after coding a 1, p0 -= T[p0 >> 8]; after a 0, p0 += T[255 - (p0 >> 8)].
const uint16_t kAdapt[256] = {
    0,    2,    5,    8,    11,   15,   20,   24,   29,   35,   41,   47,   53,   60,   67,   74,   82,   89,   97,
    106,  114,  123,  132,  141,  150,  160,  170,  180,  190,  201,  211,  222,  233,  244,  256,  267,  279,  291,
    303,  315,  327,  340,  353,  366,  379,  392,  405,  419,  433,  447,  461,  475,  489,  504,  518,  533,  548,
    563,  578,  593,  609,  624,  640,  656,  672,  688,  705,  721,  738,  754,  771,  788,  805,  822,  840,  857,
    875,  892,  910,  928,  946,  964,  983,  1001, 1020, 1038, 1057, 1076, 1095, 1114, 1133, 1153, 1172, 1192, 1211,
    1231, 1251, 1271, 1291, 1311, 1332, 1352, 1373, 1393, 1414, 1435, 1456, 1477, 1498, 1520, 1541, 1562, 1584, 1606,
    1628, 1649, 1671, 1694, 1716, 1738, 1760, 1783, 1806, 1828, 1851, 1874, 1897, 1920, 1935, 1942, 1949, 1955, 1961,
    1968, 1974, 1980, 1985, 1991, 1996, 2001, 2006, 2011, 2016, 2021, 2025, 2029, 2033, 2037, 2040, 2044, 2047, 2050,
    2053, 2056, 2058, 2061, 2063, 2065, 2066, 2068, 2069, 2070, 2071, 2072, 2072, 2072, 2072, 2072, 2072, 2071, 2070,
    2069, 2068, 2066, 2065, 2063, 2060, 2058, 2055, 2052, 2049, 2045, 2042, 2038, 2033, 2029, 2024, 2019, 2013, 2008,
    2002, 1996, 1989, 1982, 1975, 1968, 1960, 1952, 1943, 1934, 1925, 1916, 1906, 1896, 1885, 1874, 1863, 1851, 1839,
    1827, 1814, 1800, 1786, 1772, 1757, 1742, 1727, 1710, 1694, 1676, 1659, 1640, 1622, 1602, 1582, 1561, 1540, 1518,
    1495, 1471, 1447, 1422, 1396, 1369, 1341, 1312, 1282, 1251, 1219, 1186, 1151, 1114, 1077, 1037, 995,  952,  906,
    857,  805,  750,  690,  625,  553,  471,  376,  255};

struct Contexts {
    int p0[kContexts];
    Contexts() {
        for (int &p : p0) p = 0x8000;
    }
    void adapt(int c, bool one) {
        int &p = p0[c];
        p = one ? p - kAdapt[p >> 8] : p + kAdapt[255 - (p >> 8)];
    }
};

// Encoder with a 16-bit interval [low, low + range) and pending opposite bits
// for intervals that straddle the midpoint (ArithCodecBase::EncodeSymbol and
// FlushEncoder, arith_codec.h:359-405, arith_codec.cpp:95-129).
struct Encoder {
    BitSink out;
    Contexts ctx;
    unsigned low = 0, range = 0xffff;
    int pending = 0;
    void emit(bool b) {
        out.put(b);
        for (; pending > 0; --pending) out.put(!b);
    }
    void code(bool one, int c) {
        const unsigned split = (range * (unsigned)ctx.p0[c]) >> 16;
        if (one) {
            low += split;
            range -= split;
        } else {
            range = split;
        }
        ctx.adapt(c, one);
        while (range <= 0x4000) {
            if (((low + range - 1) ^ low) >= 0x8000) {
                low ^= 0x4000;
                ++pending;
            } else {
                emit(low & 0x8000);
            }
            low = (low << 1) & 0xffff;
            range <<= 1;
        }
    }
    // unary-interleaved binarisation of value + 1 (EncodeUInt): a stop
    // decision in context bin (bin1, rising to max_bin) before each of the
    // value's bits below its top bit, which go to context max_bin + 1
    void code_uint(unsigned v, int bin1, int max_bin) {
        const unsigned long long w = (unsigned long long)v + 1;
        int top = 0;
        while ((w >> (top + 1)) != 0) ++top;
        int bin = bin1;
        code(top == 0, bin);
        for (int i = top - 1; i >= 0; --i) {
            code((w >> i) & 1, max_bin + 1);
            if (bin < max_bin) ++bin;
            code(i == 0, bin);
        }
    }
    void code_sint(int v, int bin1, int max_bin) {
        code_uint(v < 0 ? 0u - (unsigned)v : (unsigned)v, bin1, max_bin);
        if (v) code(v < 0, max_bin + 2);
    }
    void finish() {
        while (((low + range - 1) ^ low) < 0x8000) {
            emit(low & 0x8000);
            low = (low << 1) & 0xffff;
            range <<= 1;
        }
        while ((low & 0x4000) && !((low + range - 1) & 0x4000)) {
            ++pending;
            low = ((low ^ 0x4000) << 1) & 0xffff;
            range <<= 1;
        }
        out.put(low & 0x4000);
        for (int i = 0; i <= pending; ++i) out.put(!(low & 0x4000));
        pending = 0;
        out.align();
    }
};

// Decoder (ArithCodecBase::InitDecoder/DecodeSymbol, arith_codec.cpp:131-150,
// arith_codec.h:228-283): bits past the data read as ones (the reference
// appends 0xff bytes).
struct Decoder {
    const uint8_t *p;
    size_t n, pos = 0;
    Contexts ctx;
    unsigned low = 0, range = 0xffff, code = 0;
    Decoder(const uint8_t *q, size_t len) : p(q), n(len) {
        for (int i = 0; i < 16; ++i) code = (code << 1) | next();
    }
    unsigned next() {
        const unsigned b = pos < 8 * n ? (p[pos >> 3] >> (7 - (pos & 7))) & 1u : 1u;
        ++pos;
        return b;
    }
    bool decode(int c) {
        const unsigned split = (range * (unsigned)ctx.p0[c]) >> 16;
        const bool one = code - low >= split;
        if (one) {
            low += split;
            range -= split;
        } else {
            range = split;
        }
        ctx.adapt(c, one);
        while (range <= 0x4000) {
            if (((low + range - 1) ^ low) >= 0x8000) {
                code ^= 0x4000;
                low ^= 0x4000;
            }
            low = (low << 1) & 0xffff;
            range <<= 1;
            code = ((code << 1) | next()) & 0xffff;
        }
        return one;
    }
    // DecodeUInt; a value that cannot fit 32 bits is refused (bad data)
    bool decode_uint(int bin1, int max_bin, unsigned &out) {
        unsigned long long v = 1;
        int bin = bin1;
        while (!decode(bin)) {
            v = (v << 1) | (unsigned)decode(max_bin + 1);
            if (v > 0x100000000ull) return false;
            if (bin < max_bin) ++bin;
        }
        out = (unsigned)(v - 1);
        return true;
    }
    bool decode_sint(int bin1, int max_bin, int &out) {
        unsigned m;
        if (!decode_uint(bin1, max_bin, m) || m > 0x7fffffffu) return false;
        out = m && decode(max_bin + 2) ? -(int)m : (int)m;
        return true;
    }
};

// ---- the MV data ------------------------------------------------------------------

// The arrays of one picture, row-major: blocks X = 4 xsb wide.
struct Field {
    int xsb, ysb, X, nrefs;
    int32_t *split, *mode;
    dme_mv *vec[2];
    int16_t *dc[3];
};

// Split prediction (SplitModeCodec::Prediction, mv_codec.cpp:98-119): the
// rounded mean of the top, top-left and left splits inside the picture, the
// left or top one on the first row or column.
int split_pred(const Field &f, int x, int y) {
    const int32_t *s = f.split;
    const int W = f.xsb;
    if (x > 0 && y > 0) return (int)(((unsigned)s[(y - 1) * W + x] + (unsigned)s[(y - 1) * W + x - 1] +
                                       (unsigned)s[y * W + x - 1] + 1u) / 3u);
    if (x > 0) return s[y * W + x - 1];
    if (y > 0) return s[(y - 1) * W + x];
    return 0;
}

// Prediction-mode prediction (PredModeCodec::Prediction, mv_codec.cpp:
// 222-253): per reference bit, the majority of the top, top-left and left
// blocks; the left or top block's mode on the first row or column.
unsigned mode_pred(const Field &f, int x, int y) {
    const int32_t *m = f.mode;
    const int X = f.X;
    if (x > 0 && y > 0) {
        const unsigned a = (unsigned)m[(y - 1) * X + x], b = (unsigned)m[(y - 1) * X + x - 1],
                       c = (unsigned)m[y * X + x - 1];
        unsigned r = ((a & 1) + (b & 1) + (c & 1)) >> 1;
        if (f.nrefs == 2) r ^= ((((a & 2) + (b & 2) + (c & 2)) >> 1) >> 1) << 1;
        return r;
    }
    if (x > 0) return (unsigned)m[x - 1];
    if (y > 0) return (unsigned)m[(y - 1) * X];
    return (unsigned)DME_INTRA;
}

// Median of the up to three available values (Median of a list,
// motion.cpp:523-556): one value, the rounded mean of two, the median of three.
// (Sums in 64 bits: the results are those of the reference's int arithmetic
// wherever that does not overflow.)
int median_list(const int *v, int n) {
    if (n == 1) return v[0];
    if (n == 2) return (int)(((long long)v[0] + v[1] + 1) >> 1);
    if (n == 3) return std::max(std::min(v[0], v[1]), std::min(std::max(v[0], v[1]), v[2]));
    return 0;
}

// a + b wrapped to 32 bits (decoding arbitrary data must not overflow)
int wrap_add(int a, int b) { return (int)(uint32_t)((uint32_t)a + (uint32_t)b); }

// Vector-element prediction of reference ref (1 or 2), component hv
// (VectorElementCodec::Prediction, mv_codec.cpp:420-460): the median of the
// top, top-left and left blocks that use the reference.
int vec_pred(const Field &f, int ref, int hv, int x, int y) {
    const int X = f.X;
    const dme_mv *v = f.vec[ref - 1];
    auto comp = [&](int i) { return hv ? v[i].y : v[i].x; };
    auto uses = [&](int i) { return (f.mode[i] & ref) != 0; };
    if (x > 0 && y > 0) {
        int n = 0, a[3];
        const int idx[3] = {(y - 1) * X + x, (y - 1) * X + x - 1, y * X + x - 1};
        for (int i : idx)
            if (uses(i)) a[n++] = comp(i);
        return n ? median_list(a, n) : 0;
    }
    if (x > 0) return uses(x - 1) ? comp(x - 1) : 0;
    if (y > 0) return uses((y - 1) * X) ? comp((y - 1) * X) : 0;
    return 0;
}

// Mean of signed values rounded as GetSMean (motion.cpp:697-722) does: half
// up for a non-negative sum; for a negative one the mean of the shifted sum
// s (1 - n) plus s.
int smean(const int *v, int n) {
    long long s = 0;
    for (int i = 0; i < n; ++i) s += v[i];
    if (s >= 0) return (int)((s + (n >> 1)) / n);
    const long long t = s - n * s;
    return (int)((t + (n >> 1)) / n + s);
}

// DC prediction of component c (DCCodec::Prediction, mv_codec.cpp:572-612):
// the mean of the intra blocks among top, top-left and left, as a short.
int dc_pred(const Field &f, int c, int x, int y) {
    const int X = f.X;
    const int16_t *d = f.dc[c];
    auto intra = [&](int i) { return f.mode[i] == DME_INTRA; };
    if (x > 0 && y > 0) {
        int n = 0, a[3];
        const int idx[3] = {(y - 1) * X + x, (y - 1) * X + x - 1, y * X + x - 1};
        for (int i : idx)
            if (intra(i)) a[n++] = d[i];
        return n ? (int16_t)smean(a, n) : 0;
    }
    if (x > 0) return intra(x - 1) ? d[x - 1] : 0;
    if (y > 0) return intra((y - 1) * X) ? d[(y - 1) * X] : 0;
    return 0;
}

// Visits the prediction units of every superblock in raster order: fn(x, y,
// step) with (x, y) the unit's top-left block and step its side in blocks.
template <class Fn>
void for_units(const Field &f, Fn fn) {
    for (int ys = 0; ys < f.ysb; ++ys)
        for (int xs = 0; xs < f.xsb; ++xs) {
            const int step = 4 >> f.split[ys * f.xsb + xs];
            for (int y = 4 * ys; y < 4 * ys + 4; y += step)
                for (int x = 4 * xs; x < 4 * xs + 4; x += step) fn(x, y, step);
        }
}

int check_info(const dme_mvdata_info *info) {
    if (!info) return set_error(DME_ERR_INVALID, "NULL info");
    if (info->xnum_sb < 1 || info->ynum_sb < 1 || info->xnum_sb > 4096 || info->ynum_sb > 4096)
        return set_error(DME_ERR_INVALID, "superblock grid out of range");
    if (info->num_refs != 1 && info->num_refs != 2) return set_error(DME_ERR_INVALID, "num_refs must be 1 or 2");
    return DME_OK;
}

int make_field(const dme_mvdata_info *info, const dme_medata *md, Field &f) {
    if (!md) return set_error(DME_ERR_INVALID, "NULL MEData");
    f.xsb = info->xnum_sb;
    f.ysb = info->ynum_sb;
    f.X = 4 * f.xsb;
    f.nrefs = info->num_refs;
    f.split = md->sb_split;
    f.mode = md->mode;
    f.vec[0] = md->vectors[0];
    f.vec[1] = md->vectors[1];
    for (int c = 0; c < 3; ++c) f.dc[c] = md->dc[c];
    if (!f.split || !f.mode || !f.vec[0] || (f.nrefs == 2 && !f.vec[1]) || !f.dc[0] || !f.dc[1] || !f.dc[2])
        return set_error(DME_ERR_INVALID, "MEData array missing (sb_split, mode, vectors, dc[3] are needed)");
    return DME_OK;
}

// The bytes of one picture's MV data.
int encode(const dme_mvdata_info *info, const dme_medata *md, std::vector<uint8_t> &out) {
    MV_TRY(check_info(info));
    Field f;
    MV_TRY(make_field(info, md, f));
    const int S = f.xsb * f.ysb, B = 16 * S;
    for (int i = 0; i < S; ++i)
        if (f.split[i] < 0 || f.split[i] > 2) return set_error(DME_ERR_INVALID, "split mode outside 0..2");
    for (int i = 0; i < B; ++i)
        if (f.mode[i] < 0 || f.mode[i] > (f.nrefs == 2 ? 3 : 1))
            return set_error(DME_ERR_INVALID, "prediction mode outside the references' modes");
    if (info->mv_precision < 0 || info->mv_precision > 3) return set_error(DME_ERR_INVALID, "mv_precision outside 0..3");
    if (info->xblen < 0 || info->yblen < 0 || info->xbsep < 0 || info->ybsep < 0)
        return set_error(DME_ERR_INVALID, "negative block parameters");
    if (info->weights_bits < 0) return set_error(DME_ERR_INVALID, "negative weight precision");
    for (int r = 0; r < f.nrefs; ++r)
        for (int i = 0; i < B; ++i)
            if (std::abs((long long)f.vec[r][i].x) > (1 << 29) || std::abs((long long)f.vec[r][i].y) > (1 << 29))
                return set_error(DME_ERR_INVALID, "vector component beyond 2^29");

    // header (MvDataByteIO::Output, mvdata_byteio.cpp:150-270)
    BitSink h;
    static const int kPresets[4][4] = {{8, 8, 4, 4}, {12, 12, 8, 8}, {16, 16, 12, 12}, {24, 24, 16, 16}};
    int pidx = 0;
    for (int i = 0; i < 4; ++i)
        if (info->xblen == kPresets[i][0] && info->yblen == kPresets[i][1] && info->xbsep == kPresets[i][2] &&
            info->ybsep == kPresets[i][3])
            pidx = i + 1;
    h.put_uint((unsigned)pidx);
    if (!pidx) {
        h.put_uint((unsigned)info->xblen);
        h.put_uint((unsigned)info->yblen);
        h.put_uint((unsigned)info->xbsep);
        h.put_uint((unsigned)info->ybsep);
    }
    h.put_uint((unsigned)info->mv_precision);
    h.put(false);  // no global motion
    h.put_uint(0);  // default frame prediction mode
    if (info->weights_bits != 1 || info->ref1_weight != 1 || (f.nrefs > 1 && info->ref2_weight != 1)) {
        h.put(true);
        h.put_uint((unsigned)info->weights_bits);
        h.put_sint(info->ref1_weight);
        if (f.nrefs > 1) h.put_sint(info->ref2_weight);
    } else {
        h.put(false);
    }
    h.align();
    out = std::move(h.bytes);

    auto element = [&](Encoder &e) {
        e.finish();
        BitSink sz;
        sz.put_uint((unsigned)e.out.bytes.size());
        sz.align();
        out.insert(out.end(), sz.bytes.begin(), sz.bytes.end());
        out.insert(out.end(), e.out.bytes.begin(), e.out.bytes.end());
    };
    {  // split modes, mod 3 from their prediction (SplitModeCodec, mv_codec.cpp:121-131)
        Encoder e;
        for (int y = 0; y < f.ysb; ++y)
            for (int x = 0; x < f.xsb; ++x) {
                int v = f.split[y * f.xsb + x] - split_pred(f, x, y);
                if (v < 0) v += 3;
                e.code_uint((unsigned)v, kSplitBin1, kSplitBin2);
            }
        element(e);
    }
    {  // prediction modes, xor their prediction bit per reference (mv_codec.cpp:255-272)
        Encoder e;
        for_units(f, [&](int x, int y, int) {
            const unsigned res = (unsigned)f.mode[y * f.X + x] ^ mode_pred(f, x, y);
            e.code(res & 1, kPmodeBit0);
            if (f.nrefs == 2) e.code(res & 2, kPmodeBit1);
        });
        element(e);
    }
    for (int ref = 1; ref <= f.nrefs; ++ref)
        for (int hv = 0; hv < 2; ++hv) {  // vector elements of units using ref (mv_codec.cpp:462-469)
            Encoder e;
            for_units(f, [&](int x, int y, int) {
                const int i = y * f.X + x;
                if (!(f.mode[i] & ref)) return;
                const dme_mv &v = f.vec[ref - 1][i];
                e.code_sint((hv ? v.y : v.x) - vec_pred(f, ref, hv, x, y), kMvBin1, kMvBin5);
            });
            element(e);
        }
    for (int c = 0; c < 3; ++c) {  // DC of intra units (mv_codec.cpp:614-620)
        Encoder e;
        for_units(f, [&](int x, int y, int) {
            const int i = y * f.X + x;
            if (f.mode[i] != DME_INTRA) return;
            e.code_sint((int)f.dc[c][i] - dc_pred(f, c, x, y), kDcBin1, kDcBin2);
        });
        element(e);
    }
    return DME_OK;
}

// Reads one picture's MV data; the arrays are written only if all of it is
// well formed.
int decode(const uint8_t *buf, size_t len, dme_mvdata_info *info, dme_medata *md, size_t *consumed) {
    if (!buf) return set_error(DME_ERR_INVALID, "NULL buffer");
    MV_TRY(check_info(info));
    Field out;
    MV_TRY(make_field(info, md, out));
    const int S = out.xsb * out.ysb, B = 16 * S;
    std::vector<int32_t> split(S, 0), mode(B, 0);
    std::vector<dme_mv> vec[2] = {std::vector<dme_mv>(B, dme_mv{0, 0}), std::vector<dme_mv>(B, dme_mv{0, 0})};
    std::vector<int16_t> dc[3] = {std::vector<int16_t>(B, 0), std::vector<int16_t>(B, 0), std::vector<int16_t>(B, 0)};
    Field f = out;
    f.split = split.data();
    f.mode = mode.data();
    f.vec[0] = vec[0].data();
    f.vec[1] = vec[1].data();
    for (int c = 0; c < 3; ++c) f.dc[c] = dc[c].data();

    BitSource h(buf, len);
    dme_mvdata_info hd = *info;
    const unsigned pidx = h.get_uint();
    static const int kPresets[4][4] = {{8, 8, 4, 4}, {12, 12, 8, 8}, {16, 16, 12, 12}, {24, 24, 16, 16}};
    if (pidx == 0) {
        hd.xblen = (int32_t)h.get_uint();
        hd.yblen = (int32_t)h.get_uint();
        hd.xbsep = (int32_t)h.get_uint();
        hd.ybsep = (int32_t)h.get_uint();
    } else if (pidx <= 4) {
        hd.xblen = kPresets[pidx - 1][0];
        hd.yblen = kPresets[pidx - 1][1];
        hd.xbsep = kPresets[pidx - 1][2];
        hd.ybsep = kPresets[pidx - 1][3];
    } else {
        return set_error(DME_ERR_INVALID, "block parameters index out of range [0-4]");
    }
    const unsigned prec = h.get_uint();
    if (!h.fail && prec > 3) return set_error(DME_ERR_INVALID, "unknown MV precision");
    hd.mv_precision = (int32_t)prec;
    if (h.get()) return set_error(DME_ERR_INVALID, "global motion parameters are not supported");
    if (h.get_uint() != 0 && !h.fail) return set_error(DME_ERR_INVALID, "non-default picture prediction mode");
    if (h.get()) {
        hd.weights_bits = (int32_t)h.get_uint();
        hd.ref1_weight = h.get_sint();
        hd.ref2_weight = out.nrefs > 1 ? h.get_sint() : 0;
    } else {
        hd.weights_bits = hd.ref1_weight = hd.ref2_weight = 1;
    }
    if (h.fail) return set_error(DME_ERR_INVALID, "MV data header truncated");
    h.align();
    size_t at = h.pos >> 3;

    // an element: its byte count, then a decoder over its data
    auto element = [&](const char *what, Decoder *&d, std::vector<Decoder> &keep) -> int {
        BitSource s(buf + at, len - at);
        const unsigned n = s.get_uint();
        s.align();
        if (s.fail) return set_error(DME_ERR_INVALID, std::string(what) + ": element size truncated");
        const size_t start = at + (s.pos >> 3);
        if (n > len - start) return set_error(DME_ERR_INVALID, std::string(what) + ": element exceeds the data");
        keep.emplace_back(buf + start, (size_t)n);
        d = &keep.back();
        at = start + n;
        return DME_OK;
    };
    std::vector<Decoder> keep;
    keep.reserve(8);
    Decoder *d = nullptr;
    bool bad = false;
    MV_TRY(element("split modes", d, keep));
    for (int y = 0; y < f.ysb; ++y)
        for (int x = 0; x < f.xsb; ++x) {
            unsigned v;
            if (!d->decode_uint(kSplitBin1, kSplitBin2, v)) return set_error(DME_ERR_INVALID, "split modes: bad code");
            split[y * f.xsb + x] = (int32_t)((v + (unsigned)split_pred(f, x, y)) % 3u);
        }
    // a unit's value reaches all its blocks (the decoder's propagation)
    auto fill = [&](auto *a, int x, int y, int step) {
        for (int j = y; j < y + step; ++j)
            for (int i = x; i < x + step; ++i) a[j * f.X + i] = a[y * f.X + x];
    };
    MV_TRY(element("prediction modes", d, keep));
    for_units(f, [&](int x, int y, int step) {
        unsigned res = d->decode(kPmodeBit0);
        if (f.nrefs == 2) res |= (unsigned)d->decode(kPmodeBit1) << 1;
        const unsigned m = mode_pred(f, x, y) ^ res;
        if (m > (f.nrefs == 2 ? 3u : 1u)) bad = true;
        mode[y * f.X + x] = (int32_t)m;
        fill(mode.data(), x, y, step);
    });
    if (bad) return set_error(DME_ERR_INVALID, "prediction modes: a mode outside the references");
    static const char *const kVecNames[2][2] = {{"reference 1 x", "reference 1 y"}, {"reference 2 x", "reference 2 y"}};
    for (int ref = 1; ref <= f.nrefs; ++ref)
        for (int hv = 0; hv < 2; ++hv) {
            MV_TRY(element(kVecNames[ref - 1][hv], d, keep));
            dme_mv *v = vec[ref - 1].data();
            for_units(f, [&](int x, int y, int step) {
                const int i = y * f.X + x;
                if (f.mode[i] & ref) {
                    int e = 0;
                    if (!d->decode_sint(kMvBin1, kMvBin5, e)) bad = true;
                    (hv ? v[i].y : v[i].x) = wrap_add(vec_pred(f, ref, hv, x, y), e);
                }
                for (int j = y; j < y + step; ++j)
                    for (int k = x; k < x + step; ++k) (hv ? v[j * f.X + k].y : v[j * f.X + k].x) = hv ? v[i].y : v[i].x;
            });
            if (bad) return set_error(DME_ERR_INVALID, std::string(kVecNames[ref - 1][hv]) + ": bad code");
        }
    static const char *const kDcNames[3] = {"Y DC", "U DC", "V DC"};
    for (int c = 0; c < 3; ++c) {
        MV_TRY(element(kDcNames[c], d, keep));
        for_units(f, [&](int x, int y, int step) {
            const int i = y * f.X + x;
            if (f.mode[i] == DME_INTRA) {
                int e = 0;
                if (!d->decode_sint(kDcBin1, kDcBin2, e)) bad = true;
                dc[c][i] = (int16_t)wrap_add(dc_pred(f, c, x, y), e);
            }
            fill(dc[c].data(), x, y, step);
        });
        if (bad) return set_error(DME_ERR_INVALID, std::string(kDcNames[c]) + ": bad code");
    }
    std::memcpy(out.split, split.data(), sizeof(int32_t) * S);
    std::memcpy(out.mode, mode.data(), sizeof(int32_t) * B);
    for (int r = 0; r < f.nrefs; ++r) std::memcpy(out.vec[r], vec[r].data(), sizeof(dme_mv) * B);
    for (int c = 0; c < 3; ++c) std::memcpy(out.dc[c], dc[c].data(), sizeof(int16_t) * B);
    *info = hd;
    if (consumed) *consumed = at;
    return DME_OK;
}

}  // namespace

extern "C" int dme_mvdata_encode(const dme_mvdata_info *info, const dme_medata *md, uint8_t *buf, size_t cap,
                                 size_t *len) try {
    if (!len) return set_error(DME_ERR_INVALID, "NULL len");
    std::vector<uint8_t> out;
    MV_TRY(encode(info, md, out));
    *len = out.size();
    if (!buf) return DME_OK;
    if (cap < out.size())
        return set_error(DME_ERR_INVALID, "buffer of " + std::to_string(cap) + " bytes; the MV data needs " +
                                              std::to_string(out.size()));
    std::memcpy(buf, out.data(), out.size());
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_mvdata_decode(const uint8_t *buf, size_t len, dme_mvdata_info *info, dme_medata *md,
                                 size_t *consumed) try {
    return decode(buf, len, info, md, consumed);
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_mvdata_encode_batch(int n, const dme_mvdata_info *info, const dme_medata *md, uint8_t *const *bufs,
                                       const size_t *caps, size_t *lens, int threads) try {
    if (n < 0 || (n > 0 && (!info || !md || !bufs || !caps || !lens)))
        return set_error(DME_ERR_INVALID, "bad batch arguments");
    const int nt = std::max(1, std::min(threads <= 0 ? (int)std::thread::hardware_concurrency() : threads, n));
    std::vector<int> rc(n, DME_OK);
    std::vector<std::string> msg(n);
    auto work = [&](int t) {
        for (int i = t; i < n; i += nt) {
            try {
                std::vector<uint8_t> out;
                rc[i] = encode(&info[i], &md[i], out);
                if (rc[i] != DME_OK) {
                    msg[i] = dme_last_error();
                    continue;
                }
                lens[i] = out.size();
                if (caps[i] < out.size()) {
                    rc[i] = DME_ERR_INVALID;
                    msg[i] = "picture " + std::to_string(i) + ": buffer of " + std::to_string(caps[i]) +
                             " bytes; the MV data needs " + std::to_string(out.size());
                    continue;
                }
                std::memcpy(bufs[i], out.data(), out.size());
            } catch (const std::bad_alloc &) {
                rc[i] = DME_ERR_NOMEM;
                msg[i] = "out of host memory";
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    for (int i = 0; i < n; ++i)
        if (rc[i] != DME_OK) return set_error(rc[i], msg[i]);
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}
