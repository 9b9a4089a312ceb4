// host_fuzz.cpp -- the host-only parts of libdme under AddressSanitizer and
// UndefinedBehaviorSanitizer: the .imt diagnostics-text reader and writer
// (untrusted input), the geometry rationalisation, the host lambda maps and
// the synthetic-picture generator.  Built by tests/sanitize/Makefile from the
// library's own sources (no GPU code) and run by tests/test_sanitize.py.
//
//   host_fuzz <file.imt> <iterations> <seed>
//
// 1. Round trip: every record of a reference-written .imt file is read and
//    written again; the bytes must be identical.
// 2. Mutations of that text (bytes replaced by digits, signs, separators or
//    arbitrary bytes, spans deleted, duplicated or padded with long digit
//    runs, truncation): every header and record the readers accept is read
//    again into arrays of its shapes and written back.  Any out-of-bounds
//    access, overflow or other undefined behaviour aborts the run.
// 3. Random parameter sets through dme_geometry_for, random vector fields
//    through dme_lambda_map_host, random sizes through the generator.
// 4. The MV-data wire format: random fields are encoded and decoded (the
//    decoded arrays must re-encode to the same bytes), then mutated and
//    truncated byte streams are decoded (untrusted input).
#include "../../include/dme.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

uint64_t g_state = 1;
uint64_t next() {  // xorshift64*
    g_state ^= g_state >> 12;
    g_state ^= g_state << 25;
    g_state ^= g_state >> 27;
    return g_state * 2685821657736338717ull;
}
int uni(int lo, int hi) { return lo + (int)(next() % (uint64_t)(hi - lo + 1)); }

struct Arrays {
    std::vector<dme_mv> vec[2];
    std::vector<dme_cost> pc[2], bi;
    std::vector<int32_t> mode, split;
    std::vector<float> intra, sbc;
    std::vector<int16_t> dc[3];
    dme_medata m;
    explicit Arrays(const dme_imt_picture &p) {
        const size_t B = (size_t)p.mv_xlen * p.mv_ylen, S = (size_t)p.sb_xlen * p.sb_ylen;
        std::memset(&m, 0, sizeof(m));
        for (int r = 0; r < 2; ++r) {
            vec[r].resize(B);
            pc[r].resize(B);
            m.vectors[r] = vec[r].data();
            m.pred_costs[r] = pc[r].data();
        }
        bi.resize(B);
        mode.resize(B);
        intra.resize(B);
        for (int c = 0; c < 3; ++c) {
            dc[c].resize(B);
            m.dc[c] = dc[c].data();
        }
        split.resize(S);
        sbc.resize(S);
        m.bipred_costs = bi.data();
        m.mode = mode.data();
        m.intra_costs = intra.data();
        m.sb_split = split.data();
        m.sb_costs = sbc.data();
    }
};

bool small_enough(const dme_imt_picture &p) {
    return (double)p.mv_xlen * p.mv_ylen <= 4e6 && (double)p.sb_xlen * p.sb_ylen <= 4e6;
}

// Reads every header/record of text the readers accept; each accepted
// record is read again into arrays and written back.  Returns the records.
int walk(const std::string &text, bool check_round_trip) {
    size_t off = 0, used = 0;
    dme_imt_sequence seq;
    if (dme_imt_read_sequence(text.data(), text.size(), &seq, &used) != DME_OK) return 0;
    if (check_round_trip) {
        size_t n = 0;
        if (dme_imt_write_sequence(&seq, nullptr, 0, &n) != DME_OK) std::abort();
        std::string b(n, '\0');
        if (dme_imt_write_sequence(&seq, &b[0], n, &n) != DME_OK) std::abort();
        if (text.compare(0, n, b) != 0) {
            std::fprintf(stderr, "sequence header differs after a round trip\n");
            std::exit(2);
        }
    }
    off = used;
    int records = 0;
    while (off < text.size()) {
        dme_imt_picture p;
        std::memset(&p, 0, sizeof(p));
        if (dme_imt_read_picture(text.data() + off, text.size() - off, &p, nullptr, &used) != DME_OK) break;
        if (used == 0 || used > text.size() - off) std::abort();
        if (!p.intra && small_enough(p)) {
            Arrays a(p);
            dme_imt_picture q = p;
            size_t used2 = 0;
            if (dme_imt_read_picture(text.data() + off, text.size() - off, &q, &a.m, &used2) != DME_OK) std::abort();
            size_t n = 0;
            if (dme_imt_write_picture(&q, &a.m, nullptr, 0, &n) == DME_OK) {
                std::string b(n, '\0');
                if (dme_imt_write_picture(&q, &a.m, &b[0], n, &n) != DME_OK) std::abort();
                // a record starts with a blank line, which the reader took
                // as the previous item's trailing whitespace
                const size_t lead = b.find_first_not_of('\n');
                if (check_round_trip && (lead == std::string::npos || text.compare(off, n - lead, b, lead, n - lead) != 0)) {
                    std::fprintf(stderr, "record %d differs after a round trip\n", records);
                    std::exit(2);
                }
            } else if (check_round_trip) {
                std::abort();
            }
        }
        off += used;
        ++records;
    }
    return records;
}

std::string mutate(const std::string &t) {
    static const char kChars[] = "0123456789-+.eE \n\t:[]>";
    std::string s = t;
    const int n = uni(1, 8);
    for (int k = 0; k < n && !s.empty(); ++k) {
        const size_t at = next() % s.size();
        switch (uni(0, 6)) {
            case 0: s[at] = kChars[next() % (sizeof(kChars) - 1)]; break;
            case 1: s[at] = (char)(next() & 0xff); break;
            case 2: s.erase(at, std::min<size_t>(s.size() - at, (size_t)uni(1, 64))); break;
            case 3: s.insert(at, s.substr(at, std::min<size_t>(s.size() - at, (size_t)uni(1, 256)))); break;
            case 4: s.insert(at, std::string((size_t)uni(1, 40), (char)('0' + uni(0, 9)))); break;
            case 5: s.insert(at, uni(0, 1) ? "-2147483649" : "99999999999999999999e99"); break;
            default: s.resize(at); break;
        }
    }
    return s;
}

void geometry_and_lambda() {
    dme_params p;
    std::memset(&p, 0, sizeof(p));
    p.xblen = uni(-4, 80);
    p.yblen = uni(-4, 80);
    p.xbsep = uni(-4, 64);
    p.ybsep = uni(-4, 64);
    p.mv_precision = uni(-1, 4);
    p.depth = uni(-2, 10);
    p.lambda = (float)uni(-5, 50) * 0.5f;
    p.fp_semantics = uni(0, 2);
    p.chroma_format = uni(0, 10) ? 0 : 1;
    p.full_search = uni(0, 3) ? 0 : 1;
    p.x_range_me = uni(-2, 140);
    p.y_range_me = uni(-2, 140);
    p.combined_me = uni(0, 2);
    const int W = uni(0, 3) ? uni(16, 1024) : uni(-8, 40000), H = uni(0, 3) ? uni(16, 768) : uni(-8, 40000);
    dme_geometry g;
    if (dme_geometry_for(&p, W, H, &g) != DME_OK) return;
    const size_t B = (size_t)g.xnum_blocks * g.ynum_blocks;
    if (B > 200000) return;
    std::vector<dme_mv> v1(B), v2(B);
    for (size_t i = 0; i < B; ++i) {
        v1[i].x = uni(-300, 300);
        v1[i].y = uni(-300, 300);
        v2[i].x = uni(-300, 300);
        v2[i].y = uni(-300, 300);
    }
    std::vector<float> lam(B);
    (void)dme_lambda_map_host(g.xnum_blocks, g.ynum_blocks, uni(1, 2), v1.data(), v2.data(), p.lambda, lam.data());
}

void synthesis() {
    const int W = 2 * uni(1, 80), H = 2 * uni(1, 60);
    std::vector<int16_t> y((size_t)W * H), r((size_t)W * H), u((size_t)(W / 2) * (H / 2)), v(u.size());
    if (dme_synth_texture(next(), W, H, uni(0, 5), y.data()) != DME_OK) return;
    dme_synth_object obj[3];
    for (auto &o : obj) {
        o.x0 = (float)uni(-20, W + 20);
        o.y0 = (float)uni(-20, H + 20);
        o.x1 = o.x0 + (float)uni(0, W);
        o.y1 = o.y0 + (float)uni(0, H);
        o.dx = (float)uni(-200, 200) * 0.25f;
        o.dy = (float)uni(-200, 200) * 0.25f;
    }
    dme_synth_motion m;
    m.zoom = 1.0f + (float)uni(-20, 20) * 0.001f;
    m.pan_x = (float)uni(-400, 400) * 0.25f;
    m.pan_y = (float)uni(-400, 400) * 0.25f;
    m.nobjects = uni(0, 3);
    m.objects = obj;
    m.noise_sigma = (float)uni(0, 8) * 0.5f;
    m.noise_seed = next();
    (void)dme_synth_warp(y.data(), W, H, &m, r.data());
    (void)dme_synth_chroma(y.data(), W, H, next(), u.data(), v.data());
}

// One random MV-data field (consistent units) through the encoder and the
// decoder, then mutations of its bytes through the decoder.
void mvdata_round(long it) {
    dme_mvdata_info info;
    std::memset(&info, 0, sizeof(info));
    info.xnum_sb = uni(1, 12);
    info.ynum_sb = uni(1, 9);
    info.num_refs = uni(1, 2);
    info.xblen = uni(4, 32);
    info.yblen = uni(4, 32);
    info.xbsep = uni(4, 32);
    info.ybsep = uni(4, 32);
    info.mv_precision = uni(0, 3);
    info.weights_bits = uni(0, 3) ? 1 : uni(0, 8);
    info.ref1_weight = uni(0, 3) ? 1 : uni(-9, 9);
    info.ref2_weight = uni(0, 3) ? 1 : uni(-9, 9);
    const int S = info.xnum_sb * info.ynum_sb, X = 4 * info.xnum_sb, B = 16 * S;
    std::vector<int32_t> split(S), mode(B);
    std::vector<dme_mv> v1(B), v2(B);
    std::vector<int16_t> dc[3] = {std::vector<int16_t>(B), std::vector<int16_t>(B), std::vector<int16_t>(B)};
    const int big = it % 4 == 0 ? 1 << 20 : 64;
    for (int s = 0; s < S; ++s) {
        split[s] = uni(0, 2);
        const int step = 4 >> split[s], xs = s % info.xnum_sb, ys = s / info.xnum_sb;
        for (int y = 4 * ys; y < 4 * ys + 4; y += step)
            for (int x = 4 * xs; x < 4 * xs + 4; x += step) {
                const int m = uni(0, info.num_refs == 2 ? 3 : 1);
                const dme_mv a{uni(-big, big), uni(-big, big)}, b{uni(-big, big), uni(-big, big)};
                const int16_t d[3] = {(int16_t)uni(-32768, 32767), (int16_t)uni(-200, 200), (int16_t)uni(-9, 9)};
                for (int j = y; j < y + step; ++j)
                    for (int i = x; i < x + step; ++i) {
                        mode[j * X + i] = m;
                        v1[j * X + i] = a;
                        v2[j * X + i] = b;
                        for (int c = 0; c < 3; ++c) dc[c][j * X + i] = d[c];
                    }
            }
    }
    dme_medata md;
    std::memset(&md, 0, sizeof(md));
    md.sb_split = split.data();
    md.mode = mode.data();
    md.vectors[0] = v1.data();
    md.vectors[1] = v2.data();
    for (int c = 0; c < 3; ++c) md.dc[c] = dc[c].data();
    size_t len = 0;
    if (dme_mvdata_encode(&info, &md, nullptr, 0, &len) != DME_OK) std::abort();
    std::vector<uint8_t> bytes(len);
    if (dme_mvdata_encode(&info, &md, bytes.data(), len, &len) != DME_OK) std::abort();
    // decode into fresh arrays and re-encode: the same bytes
    std::vector<int32_t> split2(S), mode2(B);
    std::vector<dme_mv> w1(B), w2(B);
    std::vector<int16_t> e[3] = {std::vector<int16_t>(B), std::vector<int16_t>(B), std::vector<int16_t>(B)};
    dme_medata out = md;
    out.sb_split = split2.data();
    out.mode = mode2.data();
    out.vectors[0] = w1.data();
    out.vectors[1] = w2.data();
    for (int c = 0; c < 3; ++c) out.dc[c] = e[c].data();
    dme_mvdata_info back;
    std::memset(&back, 0, sizeof(back));
    back.xnum_sb = info.xnum_sb;
    back.ynum_sb = info.ynum_sb;
    back.num_refs = info.num_refs;
    size_t used = 0;
    if (dme_mvdata_decode(bytes.data(), bytes.size(), &back, &out, &used) != DME_OK || used != bytes.size())
        std::abort();
    size_t len2 = 0;
    std::vector<uint8_t> again(len + 1);
    if (dme_mvdata_encode(&back, &out, again.data(), again.size(), &len2) != DME_OK || len2 != len ||
        std::memcmp(again.data(), bytes.data(), len) != 0)
        std::abort();
    // mutated and truncated streams: refused or decoded, never out of bounds
    for (int k = 0; k < 8; ++k) {
        std::vector<uint8_t> m = bytes;
        const int nmut = uni(1, 6);
        for (int q = 0; q < nmut && !m.empty(); ++q) m[uni(0, (int)m.size() - 1)] = (uint8_t)uni(0, 255);
        if (k & 1) m.resize(uni(0, (int)m.size()));
        dme_mvdata_info mi = back;
        (void)dme_mvdata_decode(m.data(), m.size(), &mi, &out, &used);
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: host_fuzz <file.imt> <iterations> <seed>\n");
        return 2;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    std::string text;
    char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    std::fclose(f);
    const long iters = std::atol(argv[2]);
    g_state = std::strtoull(argv[3], nullptr, 10) | 1;

    const int records = walk(text, true);
    if (records < 2) {
        std::fprintf(stderr, "the .imt file holds %d records\n", records);
        return 2;
    }
    long accepted = 0;
    for (long i = 0; i < iters; ++i) {
        accepted += walk(mutate(text), false);
        geometry_and_lambda();
        if (i % 8 == 0) synthesis();
        mvdata_round(i);
    }
    std::printf("host_fuzz: %d records round-tripped; %ld mutated texts, %ld records accepted; %ld MV-data fields "
                "round-tripped\n", records, iters, accepted, iters);
    return 0;
}
