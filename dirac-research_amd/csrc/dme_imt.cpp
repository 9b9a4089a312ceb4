// dme_imt.cpp -- the .imt diagnostics text of the encoder's MEData: the
// writer behind WriteDiagnosticsHeader / WriteDiagnosticsData
// (encoder/encmain.cpp:165-330, fed by DiracEncoder::GetInstrumentationData,
// libdirac_encoder/dirac_encoder.cpp:338-410) and the reader behind
// ReadInstrumentationHeader (util/instrumentation/instrmain.cpp:84-118) and
// ProcessSequence::AddPictureEntry (util/instrumentation/process_sequence.cpp:
// 119-210).  Host code, no GPU.
//
// Number formatting follows libstdc++'s num_put for the default stream
// state: ints in decimal, floats promoted to double and printed as "%.6g"
// (std::to_chars with chars_format::general and precision 6 is that
// conversion, locale-free).  Parsing follows num_get: an optional sign,
// digits, an optional fraction and exponent for floats, converted by
// from_chars (ints) or strtof in the "C" locale (floats); an int or short out of
// range, a float beyond the float range, or a token that is not a number, is
// an error.
#include <locale.h>
#include <stdlib.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <system_error>
#include <vector>

#include "dme_internal.h"

namespace {

using dme::set_error;

// ---- writer ------------------------------------------------------------------

// "%.6g" of a float whose result has the fixed form (decimal exponent X of
// the value rounded to six digits in [-4, 5]) and whose binary exponent keeps
// the scaled value below 2^54, computed exactly in integers: v = m * 2^e, and
// N = m * 10^k * 2^e rounded half to even (printf's rounding of the exact
// binary value) for the k = 5 - X that puts N in [10^5, 10^6).  Returns the
// length, or 0 when v is outside that domain (zero, subnormal, |v| >= 10^6 or
// below about 10^-4, inf, nan), for the general conversion to handle.
int g6_fixed(float v, char *out) {
    uint32_t bits;
    std::memcpy(&bits, &v, 4);
    const int be = (int)((bits >> 23) & 0xff);
    if (be == 0 || be == 255) return 0;
    const uint64_t m = (bits & 0x7fffff) | 0x800000;
    const int sh = 150 - be;  // v = m / 2^sh
    if (sh < 1 || sh > 40) return 0;
    static const uint64_t p10[11] = {1ull, 10ull, 100ull, 1000ull, 10000ull, 100000ull, 1000000ull, 10000000ull,
                                     100000000ull, 1000000000ull, 10000000000ull};
    auto scaled = [&](int k) {  // round_half_even(m * 10^k / 2^sh)
        const uint64_t P = m * p10[k];
        uint64_t q = P >> sh;
        const uint64_t r = P & ((1ull << sh) - 1), half = 1ull << (sh - 1);
        if (r > half || (r == half && (q & 1))) ++q;
        return q;
    };
    // a first k from the binary exponent (log10(2) ~ 0.30103), then adjust
    int k = 5 - (int)std::floor((23 - sh) * 0.30103);
    uint64_t N = 0;
    for (int it = 0; it < 4; ++it) {
        if (k < 0 || k > 10) return 0;
        N = scaled(k);
        if (N < 100000ull)
            ++k;
        else if (N >= 1000000ull)
            --k;
        else
            break;
    }
    if (N < 100000ull || N >= 1000000ull || k > 9) return 0;  // X < -4 or not settled
    const int X = 5 - k;                                      // in [-4, 5]
    char d[6];
    for (int i = 5; i >= 0; --i, N /= 10) d[i] = (char)('0' + N % 10);
    int nd = 6;
    while (nd > 1 && nd > X + 1 && d[nd - 1] == '0') --nd;  // strip fraction zeros
    int n = 0;
    if (bits >> 31) out[n++] = '-';
    if (X >= 0) {
        for (int i = 0; i <= X; ++i) out[n++] = d[i];
        if (nd > X + 1) {
            out[n++] = '.';
            for (int i = X + 1; i < nd; ++i) out[n++] = d[i];
        }
    } else {
        out[n++] = '0';
        out[n++] = '.';
        for (int i = 0; i < -X - 1; ++i) out[n++] = '0';
        for (int i = 0; i < nd; ++i) out[n++] = d[i];
    }
    return n;
}

struct Out {
    std::string s;
    void put(const char *p) { s += p; }
    void nl() { s += '\n'; }
    void i(long long v) {
        char b[24];
        auto r = std::to_chars(b, b + sizeof b, v);
        s.append(b, r.ptr);
    }
    void f(float v) {
        char b[40];
        const int n = g6_fixed(v, b);
        if (n) {
            s.append(b, n);
            return;
        }
        auto r = std::to_chars(b, b + sizeof b, (double)v, std::chars_format::general, 6);
        s.append(b, r.ptr);
    }
};

// TwoDArray operator<< (libdirac_common/arrays.h:564-578): each value then a
// space, a newline after each row.
template <class F>
void rows(Out &o, int Y, int X, F value) {
    for (int y = 0; y < Y; ++y) {
        for (int x = 0; x < X; ++x) {
            value((size_t)y * X + x);
            o.s += ' ';
        }
        o.nl();
    }
}

void cost_rows(Out &o, int Y, int X, const dme_cost *c) {
    // MvCostData operator<< (motion.cpp:381-386): "SAD mvcost".
    rows(o, Y, X, [&](size_t k) {
        o.f(c[k].SAD);
        o.s += ' ';
        o.f(c[k].mvcost);
    });
}

int finish(const std::string &s, char *buf, size_t cap, size_t *len) {
    if (!len) return set_error(DME_ERR_INVALID, "imt: len is NULL");
    *len = s.size();
    if (!buf) return DME_OK;
    if (cap < s.size())
        return set_error(DME_ERR_INVALID, "imt: buffer of " + std::to_string(cap) + " bytes, the text needs " +
                                              std::to_string(s.size()));
    std::memcpy(buf, s.data(), s.size());
    return DME_OK;
}

// ---- reader ------------------------------------------------------------------

locale_t c_locale() {
    static const locale_t loc = newlocale(LC_ALL_MASK, "C", (locale_t)0);
    return loc;
}

bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

struct In {
    const char *p, *end, *begin;
    std::string err;

    void ws() {
        while (p < end && is_space(*p)) ++p;
    }
    bool fail(const std::string &what) {
        if (err.empty()) err = "imt: " + what + " at offset " + std::to_string(p - begin);
        return false;
    }
    // The next whitespace-delimited token.
    bool token(const char *&b, const char *&e) {
        ws();
        b = p;
        while (p < end && !is_space(*p)) ++p;
        e = p;
        return b != e;
    }
    bool lit(const char *s) {
        const size_t n = std::strlen(s);
        if ((size_t)(end - p) < n || std::memcmp(p, s, n) != 0) return fail(std::string("expected \"") + s + "\"");
        p += n;
        return true;
    }
    // Digits of an integer with optional sign, directly at p (no whitespace
    // skip; used inside "[frame:N]").
    bool int_at(long long &v, const char *what) {
        const char *b = p;
        if (p < end && (*p == '+' || *p == '-')) ++p;
        const char *d = p;
        while (p < end && *p >= '0' && *p <= '9') ++p;
        if (p == d) {
            p = b;
            return fail(std::string("expected an integer for ") + what);
        }
        const char *s = b + (*b == '+');
        auto r = std::from_chars(s, p, v);
        if (r.ec != std::errc() || r.ptr != p) {
            p = b;
            return fail(std::string("integer out of range for ") + what);
        }
        return true;
    }
    bool integer(long long lo, long long hi, long long &v, const char *what) {
        ws();
        const char *b = p;
        if (!int_at(v, what)) return false;
        if (p < end && !is_space(*p)) {
            p = b;
            return fail(std::string("expected an integer for ") + what);
        }
        if (v < lo || v > hi) {
            p = b;
            return fail(std::string("value out of range for ") + what);
        }
        return true;
    }
    bool i32(int32_t &v, const char *what) {
        long long t;
        if (!integer(INT32_MIN, INT32_MAX, t, what)) return false;
        v = (int32_t)t;
        return true;
    }
    bool i16(int16_t &v, const char *what) {
        long long t;
        if (!integer(INT16_MIN, INT16_MAX, t, what)) return false;
        v = (int16_t)t;
        return true;
    }
    // [sign] digits [. digits] [(e|E) [sign] digits], at least one mantissa
    // digit, converted by strtof in the "C" locale as num_get does; a value
    // beyond the float range is an error (num_get sets failbit there).
    bool f32(float &v, const char *what) {
        const char *b, *e;
        if (!token(b, e)) return fail(std::string("expected a number for ") + what);
        const char *q = b;
        if (*q == '+' || *q == '-') ++q;
        int digits = 0;
        while (q < e && *q >= '0' && *q <= '9') ++q, ++digits;
        if (q < e && *q == '.') {
            ++q;
            while (q < e && *q >= '0' && *q <= '9') ++q, ++digits;
        }
        if (digits && q < e && (*q == 'e' || *q == 'E')) {
            ++q;
            if (q < e && (*q == '+' || *q == '-')) ++q;
            const char *x = q;
            while (q < e && *q >= '0' && *q <= '9') ++q;
            if (q == x) digits = 0;
        }
        if (!digits || q != e) {
            p = b;
            return fail(std::string("expected a number for ") + what);
        }
        const std::string tok(b, e);
        char *stop = nullptr;
        const float t = strtof_l(tok.c_str(), &stop, c_locale());
        if (stop != tok.c_str() + tok.size() || std::isinf(t)) {
            p = b;
            return fail(std::string("number out of float range for ") + what);
        }
        v = t;
        return true;
    }
    // Skips the rest of the current line and its newline.
    void line() {
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
    }
};

template <class T, class F>
bool read_rows(In &in, int Y, int X, T *dst, F one) {
    T tmp;
    for (size_t k = 0, n = (size_t)Y * X; k < n; ++k)
        if (!one(dst ? dst[k] : tmp)) return false;
    return true;
}

bool read_costs(In &in, int Y, int X, dme_cost *dst, const char *what) {
    return read_rows(in, Y, X, dst, [&](dme_cost &c) { return in.f32(c.SAD, what) && in.f32(c.mvcost, what); });
}

int read_error(In &in) { return set_error(DME_ERR_INVALID, in.err.empty() ? "imt: malformed text" : in.err); }

bool dims_ok(const dme_imt_picture &p) {
    return p.sb_ylen >= 0 && p.sb_xlen >= 0 && p.mv_ylen >= 0 && p.mv_xlen >= 0 &&
           (long long)p.mv_ylen * p.mv_xlen <= (1LL << 28) && (long long)p.sb_ylen * p.sb_xlen <= (1LL << 28);
}

}  // namespace

static int imt_write_sequence(const dme_imt_sequence *seq, char *buf, size_t cap, size_t *len) {
    if (!seq) return set_error(DME_ERR_INVALID, "imt: seq is NULL");
    Out o;
    const int32_t v[10] = {seq->chroma,         seq->width,          seq->height,         seq->source_sampling,
                           seq->topfieldfirst,  seq->frame_rate_num, seq->frame_rate_den, seq->pix_asr_num,
                           seq->pix_asr_den,    seq->picture_coding_mode};
    for (int32_t x : v) {
        o.i(x);
        o.nl();
    }
    return finish(o.s, buf, cap, len);
}

static int imt_write_picture(const dme_imt_picture *pic, const dme_medata *me, char *buf, size_t cap,
                                     size_t *len) {
    if (!pic) return set_error(DME_ERR_INVALID, "imt: pic is NULL");
    Out o;
    o.nl();
    o.put("[frame:");
    o.i(pic->pic_num);
    o.put("]");
    if (pic->intra) {
        o.put(">intra");
        o.nl();
        return finish(o.s, buf, cap, len);
    }
    const dme_imt_picture &p = *pic;
    const int nr = p.num_refs;
    if (nr < 1 || nr > 2) return set_error(DME_ERR_INVALID, "imt: num_refs must be 1 or 2");
    if (!dims_ok(p) || p.sb_ylen < 1 || p.sb_xlen < 1 || p.mv_ylen < 1 || p.mv_xlen < 1)
        return set_error(DME_ERR_INVALID, "imt: bad array lengths");
    if (!me || !me->sb_split || !me->sb_costs || !me->mode || !me->intra_costs || !me->dc[0] || !me->dc[1] ||
        !me->dc[2] || (nr > 1 && !me->bipred_costs))
        return set_error(DME_ERR_INVALID, "imt: a MEData array the record needs is NULL");
    for (int r = 0; r < nr; ++r)
        if (!me->vectors[r] || !me->pred_costs[r])
            return set_error(DME_ERR_INVALID, "imt: vectors/pred_costs of a reference are NULL");
    const int SY = p.sb_ylen, SX = p.sb_xlen, BY = p.mv_ylen, BX = p.mv_xlen;

    // First line (encmain.cpp:214-224).
    o.put(">mo_comp");
    o.nl();
    o.nl();
    o.i(nr);
    o.s += ' ';
    for (int r = 0; r < nr; ++r) {
        o.i(p.refs[r]);
        o.s += ' ';
    }
    o.i(p.ybsep);
    o.s += ' ';
    o.i(p.xbsep);
    o.s += ' ';
    o.i(SY);
    o.s += ' ';
    o.i(SX);
    o.s += ' ';
    o.i(BY);
    o.s += ' ';
    o.i(BX);

    // The MEData (motion.cpp:452-485; encmain.cpp:224-323 writes the same).
    o.nl();
    o.nl();
    rows(o, SY, SX, [&](size_t k) { o.i(me->sb_split[k]); });
    o.nl();
    rows(o, SY, SX, [&](size_t k) { o.f(me->sb_costs[k]); });
    o.nl();
    rows(o, BY, BX, [&](size_t k) { o.i(me->mode[k]); });
    o.nl();
    rows(o, BY, BX, [&](size_t k) { o.f(me->intra_costs[k]); });
    o.nl();
    if (nr > 1) cost_rows(o, BY, BX, me->bipred_costs);
    for (int c = 0; c < 3; ++c) {
        o.nl();
        rows(o, BY, BX, [&](size_t k) { o.i(me->dc[c][k]); });
    }
    for (int r = 0; r < nr; ++r) {
        o.nl();
        // MotionVector operator<< (motion.h:161-167): "x y".
        rows(o, BY, BX, [&](size_t k) {
            o.i(me->vectors[r][k].x);
            o.s += ' ';
            o.i(me->vectors[r][k].y);
        });
        o.nl();
        cost_rows(o, BY, BX, me->pred_costs[r]);
        o.nl();
    }
    return finish(o.s, buf, cap, len);
}

static int imt_read_sequence(const char *text, size_t len, dme_imt_sequence *seq, size_t *consumed) {
    if (!text || !seq || !consumed) return set_error(DME_ERR_INVALID, "imt: NULL argument");
    In in{text, text + len, text, {}};
    dme_imt_sequence s;
    int32_t tff;
    if (!in.i32(s.chroma, "chroma") || !in.i32(s.width, "width") || !in.i32(s.height, "height") ||
        !in.i32(s.source_sampling, "source_sampling") || !in.i32(tff, "topfieldfirst") ||
        !in.i32(s.frame_rate_num, "frame rate") || !in.i32(s.frame_rate_den, "frame rate") ||
        !in.i32(s.pix_asr_num, "pixel aspect ratio") || !in.i32(s.pix_asr_den, "pixel aspect ratio") ||
        !in.i32(s.picture_coding_mode, "picture_coding_mode"))
        return read_error(in);
    if (tff != 0 && tff != 1) return set_error(DME_ERR_INVALID, "imt: topfieldfirst must be 0 or 1");
    s.topfieldfirst = tff;
    in.ws();
    *seq = s;
    *consumed = (size_t)(in.p - text);
    return DME_OK;
}

static int imt_read_picture(const char *text, size_t len, dme_imt_picture *pic, dme_medata *me,
                                    size_t *consumed) {
    if (!text || !pic || !consumed) return set_error(DME_ERR_INVALID, "imt: NULL argument");
    In in{text, text + len, text, {}};
    dme_imt_picture p;
    std::memset(&p, 0, sizeof p);
    long long num;
    in.ws();
    if (!in.lit("[frame:") || !in.int_at(num, "the picture number") || !in.lit("]>")) return read_error(in);
    if (num < INT32_MIN || num > INT32_MAX) return set_error(DME_ERR_INVALID, "imt: picture number out of range");
    p.pic_num = (int32_t)num;
    const char *b, *e;
    if (!in.token(b, e)) return set_error(DME_ERR_INVALID, "imt: missing picture sort");
    const std::string sort(b, e);
    if (sort == "intra") {
        p.intra = 1;
        in.ws();
        *pic = p;
        *consumed = (size_t)(in.p - text);
        return DME_OK;
    }
    if (sort != "mo_comp") return set_error(DME_ERR_INVALID, "imt: unknown picture sort \"" + sort + "\"");
    if (!in.i32(p.num_refs, "num_refs")) return read_error(in);
    if (p.num_refs < 1 || p.num_refs > 2) return set_error(DME_ERR_INVALID, "imt: num_refs must be 1 or 2");
    for (int r = 0; r < p.num_refs; ++r)
        if (!in.i32(p.refs[r], "a reference number")) return read_error(in);
    if (!in.i32(p.ybsep, "ybsep") || !in.i32(p.xbsep, "xbsep") || !in.i32(p.sb_ylen, "sb_ylen") ||
        !in.i32(p.sb_xlen, "sb_xlen") || !in.i32(p.mv_ylen, "mv_ylen") || !in.i32(p.mv_xlen, "mv_xlen"))
        return read_error(in);
    if (!dims_ok(p)) return set_error(DME_ERR_INVALID, "imt: bad array lengths");
    if (me && (pic->sb_ylen != p.sb_ylen || pic->sb_xlen != p.sb_xlen || pic->mv_ylen != p.mv_ylen ||
               pic->mv_xlen != p.mv_xlen))
        return set_error(DME_ERR_INVALID, "imt: the record's arrays are " + std::to_string(p.sb_ylen) + "x" +
                                              std::to_string(p.sb_xlen) + " / " + std::to_string(p.mv_ylen) + "x" +
                                              std::to_string(p.mv_xlen) + ", the caller's differ");
    in.line();  // operator>>(MEData) first skips the rest of the line
    {
        // every number takes a character and a separator: refuse a record
        // whose arrays cannot fit the text before allocating for them
        const size_t S = (size_t)p.sb_ylen * p.sb_xlen, B = (size_t)p.mv_ylen * p.mv_xlen;
        const size_t values = 2 * S + B * (2 + (p.num_refs > 1 ? 2 : 0) + 3 + 4 * (size_t)p.num_refs);
        if (values && 2 * values - 1 > (size_t)(in.end - in.p))
            return set_error(DME_ERR_INVALID, "imt: the text is shorter than the record's arrays");
    }

    // Parse everything into scratch first so that a malformed record leaves
    // the caller's arrays untouched.
    const int SY = p.sb_ylen, SX = p.sb_xlen, BY = p.mv_ylen, BX = p.mv_xlen, nr = p.num_refs;
    const size_t S = (size_t)SY * SX, B = (size_t)BY * BX;
    const bool keep = me != nullptr;
    std::vector<int32_t> split(keep ? S : 0), mode(keep ? B : 0);
    std::vector<float> sbc(keep ? S : 0), intra(keep ? B : 0);
    std::vector<dme_cost> bip(keep && nr > 1 ? B : 0), pc[2];
    std::vector<int16_t> dc[3];
    std::vector<dme_mv> mv[2];
    for (int c = 0; c < 3; ++c) dc[c].resize(keep ? B : 0);
    for (int r = 0; r < nr; ++r) {
        mv[r].resize(keep ? B : 0);
        pc[r].resize(keep ? B : 0);
    }
    auto at = [&](auto &v) { return v.empty() ? nullptr : v.data(); };
    bool ok = read_rows(in, SY, SX, at(split), [&](int32_t &v) { return in.i32(v, "SBSplit"); }) &&
              read_rows(in, SY, SX, at(sbc), [&](float &v) { return in.f32(v, "SBCosts"); }) &&
              read_rows(in, BY, BX, at(mode), [&](int32_t &v) { return in.i32(v, "Mode"); }) &&
              read_rows(in, BY, BX, at(intra), [&](float &v) { return in.f32(v, "IntraCosts"); }) &&
              (nr < 2 || read_costs(in, BY, BX, at(bip), "BiPredCosts"));
    static const char *dcn[3] = {"DC(Y_COMP)", "DC(U_COMP)", "DC(V_COMP)"};
    for (int c = 0; ok && c < 3; ++c)
        ok = read_rows(in, BY, BX, at(dc[c]), [&](int16_t &v) { return in.i16(v, dcn[c]); });
    for (int r = 0; ok && r < nr; ++r)
        ok = read_rows(in, BY, BX, at(mv[r]),
                       [&](dme_mv &v) { return in.i32(v.x, "Vectors") && in.i32(v.y, "Vectors"); }) &&
             read_costs(in, BY, BX, at(pc[r]), "PredCosts");
    if (!ok) return read_error(in);
    in.ws();

    if (keep) {
        auto put = [](auto *dst, const auto &src) {
            if (dst) std::copy(src.begin(), src.end(), dst);
        };
        put(me->sb_split, split);
        put(me->sb_costs, sbc);
        put(me->mode, mode);
        put(me->intra_costs, intra);
        auto put_costs = [](dme_cost *dst, const std::vector<dme_cost> &src) {
            if (!dst) return;
            for (size_t k = 0; k < src.size(); ++k) {
                dst[k].SAD = src[k].SAD;
                dst[k].mvcost = src[k].mvcost;
            }
        };
        if (nr > 1) put_costs(me->bipred_costs, bip);
        for (int c = 0; c < 3; ++c) put(me->dc[c], dc[c]);
        for (int r = 0; r < nr; ++r) {
            put(me->vectors[r], mv[r]);
            put_costs(me->pred_costs[r], pc[r]);
        }
    }
    *pic = p;
    *consumed = (size_t)(in.p - text);
    return DME_OK;
}

// The C entry points: no C++ exception crosses them.
template <class F>
static int guarded(F f) {
    try {
        return f();
    } catch (...) {
        return dme::exception_status();
    }
}

extern "C" int dme_imt_write_sequence(const dme_imt_sequence *seq, char *buf, size_t cap, size_t *len) {
    return guarded([&] { return imt_write_sequence(seq, buf, cap, len); });
}

extern "C" int dme_imt_write_picture(const dme_imt_picture *pic, const dme_medata *me, char *buf, size_t cap,
                                     size_t *len) {
    return guarded([&] { return imt_write_picture(pic, me, buf, cap, len); });
}

extern "C" int dme_imt_read_sequence(const char *text, size_t len, dme_imt_sequence *seq, size_t *consumed) {
    return guarded([&] { return imt_read_sequence(text, len, seq, consumed); });
}

extern "C" int dme_imt_read_picture(const char *text, size_t len, dme_imt_picture *pic, dme_medata *me,
                                    size_t *consumed) {
    return guarded([&] { return imt_read_picture(text, len, pic, me, consumed); });
}
